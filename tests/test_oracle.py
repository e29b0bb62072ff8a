"""Pins of the CPU oracle (no GPU): brute-force DFTs, numpy's and scipy's pocketfft, the
reference's own known answers (tests/test_dct.cpp, test_unwrap.cpp, test_field.cpp,
test_odt.cpp, examples/quickstart.cpp), closed forms and invariants."""

import math

import numpy as np
import pytest
import scipy.fft

from conftest import contrast_error, rel_l2

GEOM = (0.532, 1.337, 1.2, 0.2)


def dft_matrix(n, sign):
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 15, 16, 24, 25, 30, 32, 48, 60, 64])
def test_fft_matches_brute_force_dft(oracle, n):
    rng = np.random.default_rng(n)
    a = (rng.standard_normal((2, 3, n)) + 1j * rng.standard_normal((2, 3, n)))
    for forward, sign in ((True, -1), (False, 1)):
        got = oracle.fft_c2c(a, axes=[2], forward=forward, dtype=np.float64)
        want = a @ dft_matrix(n, sign).T
        assert rel_l2(got, want) < 1e-13


@pytest.mark.parametrize("shape", [(8, 16, 32), (6, 10, 12), (4, 64, 128), (3, 96, 80)])
def test_fft_float32_close_to_pocketfft(oracle, shape):
    rng = np.random.default_rng(1)
    a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    want = np.fft.fftn(a.astype(np.complex128), axes=(0, 1, 2))
    got = oracle.fft_c2c(a, axes=[0, 1, 2], forward=True, dtype=np.float32)
    assert rel_l2(got, want) < 1e-6
    inv = oracle.fft_c2c(got, axes=[0, 1, 2], forward=False, scale=1.0 / a.size,
                         dtype=np.float32)
    assert rel_l2(inv, a) < 1e-6


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 1e-6)])
@pytest.mark.parametrize("shape", [(4, 6, 8), (8, 12, 10), (16, 32, 64)])
def test_real_transforms_match_numpy(oracle, shape, dtype, tol):
    rng = np.random.default_rng(2)
    x = rng.standard_normal(shape)
    h = oracle.rfft3(x, dtype=dtype)
    assert rel_l2(h, np.fft.rfftn(x)) < tol
    back = oracle.irfft3(h, shape[2], scale=1.0 / x.size, dtype=dtype)
    assert rel_l2(back, x) < tol
    # c2r of a spectrum that is not Hermitian on its kx = 0 plane keeps the real part of the
    # inverse, as pocketfft's and numpy's do.
    spec = np.fft.rfftn(x) + 0.1j * rng.standard_normal(h.shape)
    got = oracle.irfft3(spec, shape[2], scale=1.0, dtype=np.float64)
    want = np.fft.irfftn(spec, s=shape, axes=(0, 1, 2)) * x.size
    assert rel_l2(got, want) < 1e-12


@pytest.mark.parametrize("axis", [1, 2])
def test_dct_matches_scipy_type_two(oracle, axis):
    # The reference pins its Makhoul DCT against pocketfft's DCT-II, which follows FFTW's
    # REDFT10 and is twice the unnormalised sum (tests/test_dct.cpp:22-37).
    rng = np.random.default_rng(7)
    a = rng.uniform(-1, 1, (3, 12, 10)).astype(np.float32)
    want = scipy.fft.dct(a.astype(np.float64), type=2, axis=axis) * 0.5
    got32 = oracle.dct2(a, axis, dtype=np.float32)
    assert np.max(np.abs(got32 - want)) < 1e-4
    got64 = oracle.dct2(a, axis, dtype=np.float64)
    assert np.max(np.abs(got64 - want)) < 1e-12


@pytest.mark.parametrize("axis", [1, 2])
def test_dct_inverse_round_trips(oracle, axis):
    rng = np.random.default_rng(7)
    a = rng.uniform(-1, 1, (2, 16, 8)).astype(np.float32)
    b = oracle.dct2(oracle.dct2(a, axis), axis, inverse=True)
    assert np.max(np.abs(b - a)) < 1e-5


def smooth_phase(nz, ny, nx):
    """Two Gaussian bumps and a tilt, ~40 rad peak to valley (tests/test_unwrap.cpp:15-32)."""
    f = np.float32
    z, y, x = np.meshgrid(np.arange(nz, dtype=f), np.arange(ny, dtype=f), np.arange(nx, dtype=f),
                          indexing="ij")
    u = (x - f(nx * 0.4)) / f(nx * 0.15)
    v = (y - f(ny * 0.5)) / f(ny * 0.2)
    w = (x - f(nx * 0.7)) / f(nx * 0.1)
    s = (y - f(ny * 0.3)) / f(ny * 0.1)
    return (f(30) * np.exp(-(u * u + v * v)) - (f(12) + z) * np.exp(-(w * w + s * s)) +
            f(0.05) * x).astype(f)


def wrap(p):
    p = np.asarray(p, np.float32)
    pi = np.float32(math.pi)
    return (p - np.float32(2) * pi * np.floor((p + pi) / (np.float32(2) * pi))).astype(np.float32)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_unwrap_recovers_smooth_phase(oracle, dtype):
    truth = smooth_phase(2, 96, 128)
    got = oracle.unwrap(wrap(truth), dtype=dtype)
    for z in range(2):
        offset = float(got[z, 0, 0]) - float(truth[z, 0, 0])
        assert abs(math.remainder(offset, 2 * math.pi)) < 1e-3
        assert np.max(np.abs(got[z] - truth[z] - offset)) < 2e-3


def test_unwrap_solves_the_neumann_poisson_problem(oracle):
    # The 5-point Neumann Laplacian of the least-squares phase equals the divergence of the
    # wrapped gradients (zero mean), the normal equations the unwrap solves.
    rng = np.random.default_rng(3)
    psi = wrap(rng.uniform(-3, 3, (1, 24, 20)) + smooth_phase(1, 24, 20))
    phi = oracle.unwrap(psi, dtype=np.float64)[0]
    p = psi[0].astype(np.float64)

    def wr(d):
        return d - 2 * np.pi * np.floor((d + np.pi) / (2 * np.pi))

    rho = np.zeros_like(p)
    rho[:, :-1] += wr(p[:, 1:] - p[:, :-1])
    rho[:, 1:] -= wr(p[:, 1:] - p[:, :-1])
    rho[:-1, :] += wr(p[1:, :] - p[:-1, :])
    rho[1:, :] -= wr(p[1:, :] - p[:-1, :])
    lap = np.zeros_like(phi)
    lap[:, :-1] += phi[:, 1:] - phi[:, :-1]
    lap[:, 1:] -= phi[:, 1:] - phi[:, :-1]
    lap[:-1, :] += phi[1:, :] - phi[:-1, :]
    lap[1:, :] -= phi[1:, :] - phi[:-1, :]
    assert np.max(np.abs(lap - (rho - rho.mean()))) < 1e-6


def test_rytov_puts_border_near_zero(oracle):
    phase = smooth_phase(1, 64, 64) + np.float32(4 * math.pi)
    ratio = (np.float32(0.8) * np.exp(1j * phase.astype(np.float64))).astype(np.complex64)
    r = oracle.rytov(ratio)
    assert np.max(np.abs(r.real - np.log(np.float32(0.8)))) < 1e-5
    assert np.max(np.abs(r.imag - (phase - np.float32(4 * math.pi)))) < 2e-3


# ---- field extraction (tests/test_field.cpp) ----

HOLO = 128
FIELD = 64


def holo_geom(qh):
    return qh.Geometry(0.532, 1.337, 1.2, 0.055)


def field_geom(qh):
    return qh.Geometry(0.532, 1.337, 1.2, 0.055 * HOLO / FIELD)


def on_grid(bins):
    return bins / (HOLO * 0.055)


def test_finds_carrier_sideband(oracle, qh):
    bg = qh.synthesize_holograms(np.ones((1, HOLO, HOLO), np.complex64), [[0, 0]], holo_geom(qh),
                                 carrier=(5.0, 5.0))
    expected = 5.0 * HOLO * 0.055
    iy, ix, y, x, _ = oracle.find_sideband(bg)
    assert (iy, ix) == (35, 35)
    assert abs(y - expected) < 0.05 and abs(x - expected) < 0.05
    s = qh.find_sideband(bg)
    assert (s.iy, s.ix) == (35, 35)
    assert abs(s.y - y) < 1e-4 and abs(s.x - x) < 1e-4


SPHERE = [(0.0, 0.3, -0.2, 2.0, 1.36)]


def spheres(qh, spec=SPHERE):
    return [qh.Sphere(*s) for s in spec]


def test_synthetic_hologram_gives_forward_field(oracle, qh):
    illum = np.array([[0, 0], [on_grid(3), on_grid(-5)]], np.float32)
    truth_holo = qh.forward_fields(spheres(qh), illum, holo_geom(qh), HOLO, "born")
    truth = qh.forward_fields(spheres(qh), illum, field_geom(qh), FIELD, "born")
    carrier = (on_grid(35), on_grid(35))
    holo = qh.synthesize_holograms(truth_holo, illum, holo_geom(qh), carrier=carrier)
    bg = qh.synthesize_holograms(np.ones((2, HOLO, HOLO), np.complex64), illum, holo_geom(qh),
                                 carrier=carrier)
    ratio = oracle.field_ratio(oracle.extract_fields(holo, 35, 35, 17.0, FIELD),
                               oracle.extract_fields(bg, 35, 35, 17.0, FIELD))
    assert np.max(np.abs(ratio - truth)) < 1e-4


def test_noisy_hologram_stays_close(oracle, qh):
    illum = np.array([[0.0, 0.0], [0.4, -0.7]], np.float32)
    truth_holo = qh.forward_fields(spheres(qh), illum, holo_geom(qh), HOLO, "born")
    truth = qh.forward_fields(spheres(qh), illum, field_geom(qh), FIELD, "born")
    holo = qh.synthesize_holograms(truth_holo, illum, holo_geom(qh), noise=0.04, seed=3)
    bg = qh.synthesize_holograms(np.ones((2, HOLO, HOLO), np.complex64), illum, holo_geom(qh),
                                 noise=0.04, seed=4)
    ratio = oracle.field_ratio(oracle.extract_fields(holo, 35, 35, 17.0, FIELD),
                               oracle.extract_fields(bg, 35, 35, 17.0, FIELD))
    inner = np.abs(ratio - truth)[:, 8:-8, 8:-8]
    assert math.sqrt(float(np.mean(inner.astype(np.float64) ** 2))) < 0.02


def test_auto_detection_matches_explicit_centre(oracle, qh):
    truth = qh.forward_fields(spheres(qh), [[0, 0]], holo_geom(qh), HOLO, "born")
    holo = qh.synthesize_holograms(truth, [[0, 0]], holo_geom(qh))
    iy, ix, *_ = oracle.find_sideband(holo)
    auto = oracle.extract_fields(holo, iy, ix, 17.0, HOLO)
    given = oracle.extract_fields(holo, 35, 35, 17.0, HOLO)
    assert np.max(np.abs(auto - given)) < 1e-6


# ---- reconstruction (tests/test_odt.cpp, examples/quickstart.cpp) ----

N = 64
BEAD = [(0.0, 0.0, 0.0, 3.0, 1.36)]


@pytest.fixture(scope="module")
def bead_fields(oracle, qh):
    geom = qh.Geometry(*GEOM)
    illum = qh.circular_scan(49, 0.9, geom)
    ratio = qh.forward_fields(spheres(qh, BEAD), illum, geom, N)
    return illum, oracle.rytov(ratio), qh.render_phantom(spheres(qh, BEAD), geom, N, N)


def measure(ri, truth):
    c = (np.arange(N) - N // 2) * GEOM[3]
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    inner = z * z + y * y + x * x < 0.36 * 9.0
    return float(ri[inner].mean()), float(np.sqrt(np.mean((ri.astype(np.float64) - truth) ** 2)))


def test_recovers_sphere_index_and_gp_helps(oracle, bead_fields):
    illum, fields, truth = bead_fields
    plain, _ = oracle.reconstruct(fields, illum, GEOM)
    gp, _ = oracle.reconstruct(fields, illum, GEOM, iterations=20)
    m0 = measure(plain, truth)
    m1 = measure(gp, truth)
    assert abs(m0[0] - 1.36) < 4e-3
    assert abs(m1[0] - 1.36) < 4e-3
    assert m1[1] < 0.7 * m0[1]


def test_quickstart_known_answer(oracle, bead_fields):
    # examples/quickstart.cpp prints "n at the centre 1.3639 (true 1.3600), k-space filled 19.0%".
    illum, fields, _ = bead_fields
    ri, filled = oracle.reconstruct(fields, illum, GEOM, iterations=20)
    assert round(float(ri[32, 32, 32]), 4) == 1.3639
    assert round(100 * filled, 1) == 19.0


def test_real_gp_matches_complex_gp(oracle, bead_fields):
    illum, fields, _ = bead_fields
    a, fa = oracle.reconstruct(fields, illum, GEOM, iterations=10, gp_transform="r2c")
    b, fb = oracle.reconstruct(fields, illum, GEOM, iterations=10, gp_transform="c2c")
    assert contrast_error(a, b, GEOM[1]) < 1e-4
    assert fa == fb


def test_tv_reduces_error_against_gp(oracle, bead_fields):
    illum, fields, truth = bead_fields
    gp, _ = oracle.reconstruct(fields, illum, GEOM, iterations=20)
    tv, _ = oracle.reconstruct(fields, illum, GEOM, iterations=20, solver="tv")
    m_gp = measure(gp, truth)
    m_tv = measure(tv, truth)
    assert abs(m_tv[0] - 1.36) < 4e-3
    assert m_tv[1] < m_gp[1]


def test_float64_statement_agrees(oracle, bead_fields):
    illum, fields, _ = bead_fields
    a, _ = oracle.reconstruct(fields, illum, GEOM, iterations=5, solver="tv")
    b, _ = oracle.reconstruct(fields.astype(np.complex128), illum, GEOM, iterations=5,
                              solver="tv", dtype=np.float64)
    assert contrast_error(a, b, GEOM[1]) < 1e-4


def test_ewald_bins_against_double_evaluation(oracle, qh):
    # Float bins against the same geometry evaluated in double: they differ only where the
    # double value sits within float rounding of a bin boundary or of the pupil edge.
    geom = qh.Geometry(*GEOM)
    illum = qh.circular_scan(25, 0.9, geom)
    kz, nuz = oracle.ewald_targets(N, N, GEOM, illum)
    wl, nm, na, px = (float(np.float32(v)) for v in GEOM)
    fm2 = (nm / wl) ** 2
    fna2 = (na / wl) ** 2
    df = 1.0 / (N * px)
    k = np.arange(N)
    sb = np.where(k < N // 2, k, k - N) * df
    ties = 0
    for a in range(len(illum)):
        fy = sb[:, None] + float(illum[a, 0])
        fx = sb[None, :] + float(illum[a, 1])
        rho2 = fy * fy + fx * fx
        ok = (rho2 <= fna2) & (rho2 < fm2)
        fiz = math.sqrt(fm2 - float(illum[a, 0]) ** 2 - float(illum[a, 1]) ** 2)
        t = (np.sqrt(np.maximum(fm2 - rho2, 0)) - fiz) / (1.0 / (N * px)) + 0.5
        kk = np.floor(t).astype(int)
        ok &= (kk >= -N // 2) & (kk < N // 2)
        want = np.where(ok, np.where(kk < 0, kk + N, kk), -1)
        near = (np.abs(t - np.round(t)) < 1e-4) | (np.abs(rho2 - fna2) < 1e-4 * fna2)
        diff = want != kz[a]
        assert not np.any(diff & ~near)
        ties += int(np.sum(diff))
    assert ties < 10


def test_mapping_counts_equal_bin_histogram(oracle, qh, bead_fields):
    # Every accepted (illumination, column) pair adds exactly one hit to its target voxel.
    illum, fields, _ = bead_fields
    kz, _ = oracle.ewald_targets(N, N, GEOM, illum)
    _, filled, _, w = oracle.reconstruct(fields, illum, GEOM, intermediates=True)
    hist = np.zeros((N, N * N), np.int64)
    cols = np.broadcast_to(np.arange(N * N), kz.reshape(len(illum), -1).shape)
    sel = kz.reshape(len(illum), -1) >= 0
    np.add.at(hist, (kz.reshape(len(illum), -1)[sel], cols[sel]), 1)
    assert np.array_equal(hist.reshape(N, N, N), w.astype(np.int64))
    assert filled == np.count_nonzero(hist) / hist.size


def test_tv_prox_step_closed_form(oracle):
    # A step along x, constant across y and z: the isotropic TV prox decouples into 1D ROF
    # problems whose solution moves each plateau of length L by lambda / L towards the other.
    b = np.zeros((4, 4, 16), np.float64)
    b[..., :8] = 0.2
    b[..., 8:] = 1.0
    lam = 0.05
    x = oracle.tv_prox(b, lam, 3000, dtype=np.float64)
    assert np.max(np.abs(x[..., :8] - (0.2 + lam / 8))) < 1e-5
    assert np.max(np.abs(x[..., 8:] - (1.0 - lam / 8))) < 1e-5
    # A constant volume has zero TV: the prox returns it; negative data are clamped at 0.
    c = np.full((3, 5, 6), 0.7)
    assert np.max(np.abs(oracle.tv_prox(c, 0.3, 20, dtype=np.float64) - c)) < 1e-12
    assert np.all(oracle.tv_prox(-c, 0.3, 20, dtype=np.float64) == 0)


def test_circular_scan_matches_library(oracle, qh):
    geom = qh.Geometry(0.532, 1.337, 1.2, 0.11)
    a = oracle.circular_scan(49, 0.9, 0.532)
    b = qh.circular_scan(49, 0.9, geom)
    assert np.array_equal(a, b)
    assert np.array_equal(a[0], [0, 0])
    assert np.allclose(np.hypot(a[1:, 0], a[1:, 1]), 0.9 / 0.532, rtol=1e-6)
