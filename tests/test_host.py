"""Host-side checks that run without a GPU: the C-ABI library loads and exports what
include/qoraholo_hip.h declares, argument validation follows the reference's rules and raises
the exceptions its bindings raise, and the host helpers (crop plan, grid, synthesiser)."""

import ctypes
import math
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "qoraholo_hip.h").read_text()
    return sorted(set(re.findall(r"\b(qh_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from qoraholo_amd import _native

    lib = _native.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(_native.symbols()) == names


def test_unknown_backend_is_rejected(qh):
    with pytest.raises(ValueError, match="backend"):
        qh.unwrap(np.zeros((8, 8), np.float32), backend="cuda")
    with pytest.raises(ValueError, match="backend"):
        qh.reconstruct(np.zeros((1, 8, 8), np.complex64), [[0, 0]], qh.Geometry(), backend="cpu")


def test_validation_errors_are_value_errors(qh):
    g = qh.Geometry(0.532, 1.337, 1.2, 0.2)
    with pytest.raises(ValueError, match="square with an even side"):
        qh.extract_fields(np.zeros((1, 10, 12), np.float32), center=(2, 2))
    with pytest.raises(ValueError, match="crop radius"):
        qh.extract_fields(np.zeros((1, 64, 64), np.float32), center=(10, 10), radius=40,
                          output_size=64)
    with pytest.raises(ValueError, match="output_size"):
        qh.extract_fields(np.zeros((1, 64, 64), np.float32), center=(10, 10), output_size=30 + 1)
    with pytest.raises(ValueError, match="even"):
        qh.unwrap(np.zeros((1, 9, 8), np.float32))
    with pytest.raises(ValueError, match="illumination frequency"):
        qh.reconstruct(np.zeros((1, 16, 16), np.complex64), [[3.0, 0.0]], g)
    with pytest.raises(ValueError, match="nz must be even"):
        qh.reconstruct(np.zeros((1, 16, 16), np.complex64), [[0.0, 0.0]], g, nz=15)
    with pytest.raises(ValueError, match="2\\^a 3\\^b 5\\^c"):
        qh.reconstruct(np.zeros((1, 14, 14), np.complex64), [[0.0, 0.0]], g)
    with pytest.raises(ValueError, match="one illumination per field"):
        qh.reconstruct(np.zeros((2, 16, 16), np.complex64), [[0.0, 0.0]], g)
    with pytest.raises(ValueError, match="unknown solver"):
        qh.reconstruct(np.zeros((1, 16, 16), np.complex64), [[0.0, 0.0]], g, solver="admm")


def test_pipeline_validation(qh):
    g = qh.Geometry(0.532, 1.337, 1.2, 0.2)
    il = qh.circular_scan(4, 0.9, g)
    with pytest.raises(ValueError, match="carrier"):
        qh.Pipeline(4, 64, None, il, g)
    with pytest.raises(ValueError, match="upload_chunks"):
        qh.Pipeline(4, 64, (16, 16), il, g, output_size=32, upload_chunks=5)
    with pytest.raises(ValueError, match="hologram_size"):
        qh.Pipeline(4, 63, (16, 16), il, g, output_size=32)


def test_crop_plan_rounds_and_defaults(qh):
    from qoraholo_amd import _crop

    p = _crop(256, (70.5, 70.5), 32.0, 128)
    assert (p.cy, p.cx, p.out) == (71, 71, 128)
    p = _crop(256, (-30.4, 4.6), 0.0, 0)
    assert (p.cy, p.cx, p.out) == (-30, 5, 256)
    assert p.radius == np.float32(math.hypot(np.float32(-30.4), np.float32(4.6))) / np.float32(3)


def test_grid_constants(qh):
    from qoraholo_amd import _grid

    g = _grid(256, 256, qh.Geometry(0.532, 1.337, 1.2, 0.11))
    f = np.float32
    fm = f(1.337) / f(0.532)
    assert g.fm2 == fm * fm
    assert g.df == f(1) / (f(256) * f(0.11))
    assert g.gain == f(4) * f(math.pi) / f(0.11)


def test_noise_stream_is_seeded_gaussian(qh):
    from qoraholo_amd.phantom import gaussian_noise

    a = gaussian_noise(0, 0.5, 200000)
    b = gaussian_noise(0, 0.5, 200000)
    c = gaussian_noise(1, 0.5, 1000)
    assert np.array_equal(a, b) and not np.array_equal(a[:1000], c)
    assert abs(float(a.mean())) < 0.005
    assert abs(float(a.std()) - 0.5) < 0.005


def test_hologram_synthesis_closed_form(qh):
    # Without a sample (ratio 1) the hologram of a normal-incidence wave and a unit reference is
    # |1 + exp(2 pi i c.r)|^2 = 2 + 2 cos(2 pi c.r).
    g = qh.Geometry(0.532, 1.337, 1.2, 0.1)
    h = qh.synthesize_holograms(np.ones((1, 32, 32), np.complex64), [[0, 0]], g, carrier=(1.0, 2.0))
    r = (np.arange(32) - 16) * np.float32(0.1)
    want = 2 + 2 * np.cos(2 * np.pi * (1.0 * r[:, None] + 2.0 * r[None, :]))
    assert np.max(np.abs(h[0] - want)) < 1e-5


def test_cell_shapes_are_nested(qh):
    from qoraholo_amd.workloads import cell_shapes

    body, nucleus, v1, v2 = cell_shapes(0)
    for s in (nucleus, v1, v2):
        # The farthest point of each sphere from the body centre stays inside the body.
        for d in np.eye(3):
            for sign in (-1, 1):
                p = np.array([s.z, s.y, s.x]) + sign * s.radius * d
                assert (p[0] / body.az) ** 2 + (p[1] / body.ay) ** 2 + (p[2] / body.ax) ** 2 < 1
    for a, b in ((nucleus, v1), (nucleus, v2), (v1, v2)):
        assert math.dist((a.z, a.y, a.x), (b.z, b.y, b.x)) > a.radius + b.radius


def test_bpm_tables_equal_oracle_bitwise(oracle, qh):
    """The library's tables (src/bpm.cpp:10-60, double on the host) are the oracle's bits."""
    g = qh.Geometry(0.532, 1.337, 1.2, 0.25)
    il = qh.circular_scan(5, 0.8, g)
    for n, nz, obliquity in ((48, 32, True), (32, 16, False), (64, 7, True)):
        got = qh.bpm_tables(n, nz, il, g, obliquity)
        want = oracle.bpm_tables(n, nz, il, (0.532, 1.337, 1.2, 0.25), obliquity)
        for a, b in zip(got[:3], want[:3]):
            assert np.array_equal(a, b)
        assert got[3] == want[3] and got[4] == want[4]


def test_bpm_argument_checks(qh):
    """src/api.cpp:141-157,180-188: rejected before any device work, as ValueError."""
    g = qh.Geometry(0.532, 1.337, 1.2, 0.2)
    il = qh.circular_scan(3, 0.9, g)
    ri = np.full((4, 32, 32), 1.337, np.float32)
    with pytest.raises(ValueError):
        qh.bpm_forward(np.full((4, 31, 31), 1.337, np.float32), il, g)
    with pytest.raises(ValueError):
        qh.bpm_forward(np.full((4, 32, 16), 1.337, np.float32), il, g)
    with pytest.raises(ValueError):
        qh.bpm_forward(ri, np.array([[3.0, 0.0]], np.float32), g)
    with pytest.raises(ValueError):
        qh.bpm_forward(ri, il, qh.Geometry(0.532, 1.337, 1.2, 0.0))
    with pytest.raises(ValueError):
        qh.bpm_forward(np.full((4, 28, 28), 1.337, np.float32), il, g)  # 28 = 4 * 7
    with pytest.raises(ValueError):
        qh.bpm_forward(ri, il, g, backend="cuda")
    fields = np.ones((3, 32, 32), np.complex64)
    with pytest.raises(ValueError):
        qh.bpm_reconstruct(fields, il, g, ri, step=0.0)
    with pytest.raises(ValueError):
        qh.bpm_reconstruct(fields, il, g, ri, tv_inner=0)
    with pytest.raises(ValueError):
        qh.bpm_reconstruct(fields, il, g, ri, iterations=-1)
    with pytest.raises(ValueError):
        qh.bpm_reconstruct(fields[:, :16, :16], il, g, ri)
    with pytest.raises(ValueError):
        qh.bpm_reconstruct(fields[:2], il, g, ri)
    with pytest.raises(ValueError):
        qh.bpm_vjp(ri - np.float32(1.337), fields[:2], il, g)
    with pytest.raises(ValueError, match="checkpoint at least -1"):
        qh.bpm_forward(ri, il, g, checkpoint=-2)


def test_bpm_plan_checkpoint_needs_a_plan(qh):
    """The resolved checkpoint interval is read from a plan; a null plan is refused before any
    device work, as ValueError."""
    import ctypes

    from qoraholo_amd import _native

    out = ctypes.c_int(-5)
    with pytest.raises(ValueError, match="null plan"):
        _native.check(_native.lib().qh_hip_bpm_plan_checkpoint(None, ctypes.byref(out)))
    assert out.value == -5


@pytest.mark.parametrize("batch", [0, 65536])
def test_tv_prox_batch_argument_checks(qh, batch):
    """The batched prox refuses batch sizes outside 1..65535 (the grid's second dimension)
    before any device work, as ValueError."""
    from qoraholo_amd import _native

    b = np.zeros((2, 4, 4, 4), np.float32)
    x = np.empty_like(b)
    with pytest.raises(ValueError, match="batch 1 to 65535"):
        _native.check(_native.lib().qh_hip_tv_prox_batch(_native.ptr(b), batch, 4, 4, 4, 0.1, 2, 1,
                                                         _native.ptr(x)))


def test_host_copy_threads_any_size_and_alignment():
    """qh_host_copy splits large copies over the library's copy threads: every size, offset
    and concurrent caller gets exactly the source bytes."""
    import threading

    from qoraholo_amd import _native

    rng = np.random.default_rng(11)
    src = rng.integers(0, 256, (48 << 20) + 777, dtype=np.uint8)
    for size, off in [(0, 0), (1, 3), (4096, 0), ((2 << 20) + 5, 1), ((3 << 20) - 1, 7),
                      ((33 << 20) + 123, 13)]:
        dst = np.zeros(size + 64, np.uint8)
        _native.check(_native.lib().qh_host_copy(dst.ctypes.data + off, src.ctypes.data + off,
                                                 size))
        assert np.array_equal(dst[off:off + size], src[off:off + size])
        assert not dst[:off].any() and not dst[off + size:].any()

    outs = [np.zeros(len(src), np.uint8) for _ in range(6)]
    errors = []

    def copy(i):
        try:
            for _ in range(3):
                _native.check(_native.lib().qh_host_copy(outs[i].ctypes.data, src.ctypes.data,
                                                         len(src)))
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    threads = [threading.Thread(target=copy, args=(i,)) for i in range(len(outs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors
    for o in outs:
        assert np.array_equal(o, src)


@pytest.mark.parametrize("count,bcount,shape", [(49, 1, (64, 64)), (7, 7, (33, 40)),
                                                 (3, 1, (1, 5)), (40, 40, (128, 128))])
def test_field_ratio_matches_oracle_bitwise(qh, oracle, count, bcount, shape):
    """The host field_ratio (src/cpu/field.cpp:62-74) on the library's threads equals the
    oracle's std::complex<float> loop bit for bit, for one background or one per field,
    including zeros, infinities and tiny denominators."""
    rng = np.random.default_rng(count)
    s = (rng.standard_normal((count,) + shape) + 1j * rng.standard_normal((count,) + shape))
    b = (rng.standard_normal((bcount,) + shape) + 1j * rng.standard_normal((bcount,) + shape))
    s = s.astype(np.complex64)
    b = b.astype(np.complex64)
    b.flat[0] = 0
    b.flat[-1] = 1e-30 + 1e-30j
    b.flat[2] = complex(np.inf, 1.0)
    b.flat[3] = 3e38 + 3e38j
    s.flat[1] = np.inf
    s.flat[4] = complex(np.nan, 2.0)
    s.flat[5] = 3e38 - 3e38j
    got = qh.field_ratio(s, b)
    want = oracle.field_ratio(s, b)
    assert got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
