"""The allocation-free TV prox (qh_hip_tv_prox_dev_noalloc: device pointers, the dual in the
library's fixed device array) against the allocating entry point qh_hip_tv_prox on the same
inputs, bit for bit, over consecutive calls whose volumes shrink (the fixed array keeps the
previous call's dual, which each call must clear), and its refusal of volumes above the limit."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_VOXELS = 2097152  # QH_TV_NOALLOC_MAX_VOXELS


def _prox_allocating(lib, native, b, lam, inner, fused):
    x = np.empty_like(b)
    nz, ny, nx = b.shape
    native.check(lib.qh_hip_tv_prox(native.ptr(b), nz, ny, nx, lam, inner, fused, native.ptr(x)))
    return x


def test_noalloc_prox_equals_allocating_prox(device):
    import torch
    from qoraholo_amd import _native

    lib = _native.lib()
    rng = np.random.default_rng(11)
    stream = torch.cuda.current_stream()
    # Consecutive calls, each volume smaller than the one before: two FGP steps per pass with
    # an odd last step, x tiles of a 384-wide row, single steps only.
    cases = [((40, 48, 56), 0.03, 5, 1), ((8, 20, 384), 0.02, 4, 1), ((24, 20, 32), 0.05, 3, 0)]
    for shape, lam, inner, fused in cases:
        print(f"noalloc prox {shape} inner {inner} fused {fused}", flush=True)
        b = (rng.standard_normal(shape) * 0.2 + 0.3).astype(np.float32)
        want = _prox_allocating(lib, _native, b, lam, inner, fused)
        bd = torch.from_numpy(b).cuda()
        xd = torch.full(shape, float("nan"), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        _native.check(lib.qh_hip_tv_prox_dev_noalloc(
            _native.ptr(bd.data_ptr()), *shape, lam, inner, fused, _native.ptr(xd.data_ptr()),
            _native.ptr(stream.cuda_stream)))
        torch.cuda.synchronize()
        got = xd.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), shape
        assert np.all(got >= 0)


def test_noalloc_prox_refuses_volumes_above_the_limit(device):
    import torch
    from qoraholo_amd import _native

    lib = _native.lib()
    bd = torch.zeros(16, device="cuda")
    for shape in [(129, 128, 128), (2, 1024, 1025)]:
        assert shape[0] * shape[1] * shape[2] > MAX_VOXELS
        with pytest.raises(ValueError, match="exceeds the fixed dual buffer"):
            _native.check(lib.qh_hip_tv_prox_dev_noalloc(
                _native.ptr(bd.data_ptr()), *shape, 0.01, 2, 1, _native.ptr(bd.data_ptr()),
                _native.ptr(0)))
