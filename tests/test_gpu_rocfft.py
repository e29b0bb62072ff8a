"""Cross-check of the library's transforms against rocFFT (through torch.fft on the device).
rocFFT is only the checker here; the product path never calls it."""

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(64, 64, 64), (32, 256, 256), (128, 96, 40)])
def test_3d_real_transform_matches_rocfft(device, shape):
    torch = pytest.importorskip("torch")
    from qoraholo_amd import _native

    rng = np.random.default_rng(4)
    x = rng.standard_normal(shape).astype(np.float32)
    h = np.empty(shape[:2] + (shape[2] // 2 + 1,), np.complex64)
    _native.check(_native.lib().qh_hip_rfft3(_native.ptr(x), *shape, _native.ptr(h)))
    ref = torch.fft.rfftn(torch.from_numpy(x).to("cuda:0")).cpu().numpy()
    assert rel_l2(h, ref) < 2e-6


@pytest.mark.parametrize("shape", [(49, 256, 256), (4, 512, 512)])
def test_2d_complex_transform_matches_rocfft(device, shape):
    torch = pytest.importorskip("torch")
    from qoraholo_amd import _native

    rng = np.random.default_rng(5)
    a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    got = a.copy()
    _native.check(_native.lib().qh_hip_fft_c2c(_native.ptr(got), *shape, 6, 1, 1.0))
    ref = torch.fft.fft2(torch.from_numpy(a).to("cuda:0")).cpu().numpy()
    assert rel_l2(got, ref) < 2e-6
