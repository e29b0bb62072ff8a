"""Pins of the oracle's beam propagation model (no GPU): the reference's own checks in
tests/test_bpm.cpp (adjoint identity, tangent model against finite differences, weak-object
limit against the analytic Born fields), closed forms of the tables of src/bpm.cpp:10-60, the
Lipschitz constant as a Rayleigh quotient, and the inversion's known behaviour."""

import numpy as np
import pytest

from conftest import rel_l2

GEOM_ADJ = (0.532, 1.337, 1.2, 0.25)  # tests/test_bpm.cpp:89
GEOM = (0.532, 1.337, 1.2, 0.2)  # tests/test_bpm.cpp:14


def pattern(nz, n, amplitude, seed):
    """Smooth deterministic volume of tests/test_bpm.cpp:41-53, in [0, amplitude)."""
    z, y, x = np.meshgrid(np.arange(nz), np.arange(n), np.arange(n), indexing="ij")
    s = np.sin(0.31 * x + 0.17 * y * seed + 0.53 * z) * np.cos(0.23 * y - 0.41 * z * seed + 0.11 * x)
    return (amplitude * 0.5 * (1.0 + s)).astype(np.float32)


def residual_pattern(count, n):
    """r of tests/test_bpm.cpp:94-97."""
    i = np.arange(count * n * n, dtype=np.float64)
    return (np.sin(0.7 * i) + 1j * np.cos(1.3 * i)).astype(np.complex64).reshape(count, n, n)


@pytest.fixture(scope="module")
def adj_case(oracle):
    il = oracle.circular_scan(3, 0.8, GEOM_ADJ[0])
    return il, pattern(16, 32, 0.05, 1), pattern(16, 32, 1.0, 2), residual_pattern(3, 32)


def inner_pair(jh, r, h, jtr):
    lhs = np.sum(np.real(np.conj(jh.astype(np.complex128)) * r.astype(np.complex128)))
    rhs = np.sum(h.astype(np.float64) * jtr.astype(np.float64))
    return lhs, rhs


def test_adjoint_identity_fp64(oracle, adj_case):
    """<J h, r> = <h, J^T r> at a non-zero x (tests/test_bpm.cpp:88-106), exact up to double
    rounding in the FP64 statement."""
    il, x, h, r = adj_case
    jh = oracle.bpm_jvp(x, h, il, GEOM_ADJ, dtype=np.float64)
    jtr = oracle.bpm_vjp(x, r, il, GEOM_ADJ, dtype=np.float64)
    lhs, rhs = inner_pair(jh, r, h, jtr)
    assert abs(lhs - rhs) / abs(rhs) < 1e-12


def test_adjoint_identity_float32(oracle, adj_case):
    """The float statement: the reference documents 2.7e-6 for its pocketfft build
    (docs/ARCHITECTURE.md:219); float rounding of the two 16-slice chains sets the level."""
    il, x, h, r = adj_case
    jh = oracle.bpm_jvp(x, h, il, GEOM_ADJ)
    jtr = oracle.bpm_vjp(x, r, il, GEOM_ADJ)
    lhs, rhs = inner_pair(jh, r, h, jtr)
    assert abs(lhs - rhs) / abs(rhs) < 5e-5


def test_float32_model_close_to_fp64(oracle, adj_case):
    il, x, h, r = adj_case
    f32 = oracle.bpm_forward(x, il, GEOM_ADJ)
    f64 = oracle.bpm_forward(x, il, GEOM_ADJ, dtype=np.float64)
    assert rel_l2(f32 - 1, f64 - 1) < 5e-6
    g32 = oracle.bpm_vjp(x, r, il, GEOM_ADJ)
    g64 = oracle.bpm_vjp(x, r, il, GEOM_ADJ, dtype=np.float64)
    assert rel_l2(g32, g64) < 1e-5


def test_jvp_against_central_differences(oracle, adj_case):
    """J h against a central difference of the forward model (tests/test_bpm.cpp:108-126;
    1.8e-4 documented at eps = 1e-3, docs/ARCHITECTURE.md:220). The difference is dominated by
    the eps^2 term, so in FP64 a smaller step brings it down quadratically."""
    il, x, h, _ = adj_case

    def fd_error(eps, dtype):
        jh = oracle.bpm_jvp(x, h, il, GEOM_ADJ, dtype=dtype).astype(np.complex128)
        xp = (x.astype(np.float64) + eps * h).astype(dtype)
        xm = (x.astype(np.float64) - eps * h).astype(dtype)
        fd = (oracle.bpm_forward(xp, il, GEOM_ADJ, dtype=dtype).astype(np.complex128)
              - oracle.bpm_forward(xm, il, GEOM_ADJ, dtype=dtype)) / (2 * eps)
        return rel_l2(fd, jh)

    e3 = fd_error(1e-3, np.float32)
    assert 1e-4 < e3 < 3e-4
    d3 = fd_error(1e-3, np.float64)
    d4 = fd_error(1e-4, np.float64)
    assert abs(d3 - e3) < 2e-5
    assert d4 < d3 / 50


def test_gradient_against_fp64_differences_of_the_loss(oracle, adj_case):
    """d/dx_i of 0.5/count sum |bpm(x) - d|^2 by central differences (step 1e-4, FP64) at 16
    seeded voxels against the gradient (scaled by count, as gradient() divides by the batch)."""
    il, x, h, _ = adj_case
    rng = np.random.default_rng(4)
    data = oracle.bpm_forward((x * 1.3).astype(np.float32), il, GEOM_ADJ).astype(np.complex128)
    data += 0.01 * (rng.standard_normal(data.shape) + 1j * rng.standard_normal(data.shape))
    x64 = x.astype(np.float64)
    grad, loss = oracle.bpm_gradient(x64, data, il, GEOM_ADJ, dtype=np.float64)
    count = il.shape[0]
    flat = rng.choice(x.size, 16, replace=False)
    eps = 1e-4
    for i in flat:
        xp = x64.copy().reshape(-1)
        xm = x64.copy().reshape(-1)
        xp[i] += eps
        xm[i] -= eps
        _, lp = oracle.bpm_gradient(xp.reshape(x.shape), data, il, GEOM_ADJ, dtype=np.float64)
        _, lm = oracle.bpm_gradient(xm.reshape(x.shape), data, il, GEOM_ADJ, dtype=np.float64)
        fd = (lp - lm) / (2 * eps)
        # gradient() returns Re(J^H r) / count; the loss carries 0.5 / count.
        assert fd == pytest.approx(grad.reshape(-1)[i], rel=1e-6, abs=1e-10)
    assert loss > 0


def test_weak_object_is_born_with_obliquity(oracle, qh):
    """tests/test_bpm.cpp:71-84 and README.md:334: with the obliquity weight the BPM of a weak
    sphere is the first Born field to 2.7-3.6% at every angle; the plain screen only at normal
    incidence."""
    from qoraholo_amd import phantom as ph

    g = qh.Geometry(*GEOM)
    s = [ph.Sphere(0.4, -0.6, 0.2, 3.0, 1.3375)]
    il = oracle.circular_scan(5, 0.9, GEOM[0])
    born = ph.forward_fields(s, il, g, 64, model="born")
    x = ph.render_phantom(s, g, 64, 64) - np.float32(GEOM[1])
    with_q = oracle.bpm_forward(x, il, GEOM)
    plain = oracle.bpm_forward(x, il, GEOM, obliquity=False)

    def err(u, a):
        return np.sqrt(np.sum(np.abs(u[a] - born[a]) ** 2) / np.sum(np.abs(born[a] - 1) ** 2))

    e_q = [err(with_q, a) for a in range(5)]
    e_p = [err(plain, a) for a in range(5)]
    assert all(0.02 < e < 0.04 for e in e_q), e_q
    assert e_p[0] < 0.04 and all(e > 0.2 for e in e_p[1:]), e_p


def test_tables_closed_form(oracle):
    """src/bpm.cpp:10-60 re-evaluated entry by entry in numpy double."""
    n, nz = 24, 10
    il = oracle.circular_scan(4, 0.9, GEOM[0])
    step, last, q, phase, lip = oracle.bpm_tables(n, nz, il, GEOM)
    lam, nm, na, d = (np.float64(np.float32(v)) for v in GEOM)
    fm, fna = nm / lam, na / lam
    df = 1.0 / (n * d)
    k = np.arange(n)
    sb = np.where(k < n // 2, k, k - n)
    nuz_min = np.sqrt(max(fm * fm - fna * fna, 0.01 * fm * fm))
    back = -(nz - 1 - nz // 2) * d
    for a in range(il.shape[0]):
        fiy, fix = np.float64(il[a, 0]), np.float64(il[a, 1])
        fiz = np.sqrt(fm * fm - fiy * fiy - fix * fix)
        fy = sb[:, None] * df + fiy
        fx = sb[None, :] * df + fix
        rho2 = fy * fy + fx * fx
        band = rho2 < fm * fm
        nuz = np.sqrt(np.where(band, fm * fm - rho2, 0.0))
        want_step = np.where(band, np.exp(2j * np.pi * (nuz - fiz) * d), 0)
        want_last = np.where(band & (rho2 <= fna * fna), np.exp(2j * np.pi * (nuz - fiz) * back), 0)
        want_q = np.where(band, fm / np.maximum(nuz, nuz_min), 1.0)
        assert np.max(np.abs(step[a] - want_step)) < 2e-7
        assert np.max(np.abs(last[a] - want_last)) < 2e-7
        assert np.max(np.abs(q[a] - want_q) / want_q) < 1e-7
    assert phase == pytest.approx(2 * np.pi / lam * d, rel=1e-7)
    q0 = np.array([q[a, 0, 0] for a in range(il.shape[0])], np.float64)
    assert lip == pytest.approx(nz * np.float64(phase) ** 2 * np.mean(q0 ** 2), rel=1e-7)
    _, _, q1, _, _ = oracle.bpm_tables(n, nz, il, GEOM, obliquity=False)
    assert np.all(q1 == 1.0)


def test_lipschitz_is_the_rayleigh_quotient_of_the_constant_volume(oracle, adj_case):
    """For a weak object every slice adds i phase q_a(0) x to the zero bin of every angle, so
    |J 1|^2 / (count |1|^2) at x = 0 is nz phase^2 mean(q_a(0)^2), the step's Lipschitz constant
    (include/qoraholo/bpm.hpp:37-39)."""
    il = adj_case[0]
    nz, n = 16, 32
    zero = np.zeros((nz, n, n))
    ones = np.ones((nz, n, n))
    j1 = oracle.bpm_jvp(zero, ones, il, GEOM_ADJ, dtype=np.float64)
    rq = np.sum(np.abs(j1) ** 2) / (il.shape[0] * np.sum(ones))
    _, _, _, _, lip = oracle.bpm_tables(n, nz, il, GEOM_ADJ)
    assert rq == pytest.approx(lip, rel=1e-6)


def test_reconstruction_lowers_the_loss_and_recovers_a_bead(oracle, qh):
    """Inversion of BPM data of a bead (the Python variant of the reference's check,
    python/tests/test_qoraholo.py:88-97, with model data): loss falls, RMSE falls from a flat
    start, batched and FISTA both descend; TV keeps n >= n_medium."""
    from qoraholo_amd import phantom as ph

    g = qh.Geometry(*GEOM)
    bead = [ph.Sphere(0.0, 0.0, 0.0, 1.6, 1.36)]
    il = oracle.circular_scan(9, 0.9, GEOM[0])
    truth = ph.render_phantom(bead, g, 24, 32)
    data = oracle.bpm_forward(truth - np.float32(GEOM[1]), il, GEOM)
    start = np.full_like(truth, np.float32(GEOM[1]))
    for batch, tv in ((0, 0.0), (3, 0.0), (0, 1e-2)):
        ri, loss = oracle.bpm_reconstruct(data, il, GEOM, start, iterations=8, batch=batch,
                                          tv_weight=tv)
        assert loss.shape == (8,) and loss[-1] < 0.2 * loss[0], (batch, tv, loss)
        assert np.sqrt(np.mean((ri - truth) ** 2)) < 0.5 * np.sqrt(np.mean((start - truth) ** 2))
        assert ri.min() >= np.float32(GEOM[1])


def exact_bead_case(oracle, qh):
    """tests/test_bpm.cpp:131-171: a 3 um, n = 1.40 bead, 64^3 at 0.2 um, 49 illuminations at
    NA 0.9, data from the exact partial-wave model, Rytov + GP (20) start."""
    from qoraholo_amd import phantom as ph

    g = qh.Geometry(*GEOM)
    bead = [ph.Sphere(0.0, 0.0, 0.0, 3.0, 1.40)]
    il = oracle.circular_scan(49, 0.9, GEOM[0])
    data = ph.forward_fields(bead, il, g, 64, model="exact")
    truth = ph.render_phantom(bead, g, 64, 64)
    initial, _ = oracle.reconstruct(oracle.rytov(data), il, GEOM, iterations=20)
    z, y, x = np.indices(truth.shape) - 32
    inner = (z * z + y * y + x * x) * GEOM[3] ** 2 < 0.36 * 3.0 ** 2
    return g, il, data, truth, initial, inner


def test_exact_model_matches_born_for_weak_sphere(qh):
    """tests/test_bpm.cpp:29-36: the partial-wave solution of a weak sphere is the first Born
    field to 2%."""
    from qoraholo_amd import phantom as ph

    g = qh.Geometry(*GEOM)
    s = [ph.Sphere(0.4, -0.6, 0.2, 3.0, 1.3375)]
    il = qh.circular_scan(5, 0.9, g)
    exact = ph.forward_fields(s, il, g, 64, model="exact")
    born = ph.forward_fields(s, il, g, 64, model="born")
    e = np.sqrt(np.sum(np.abs(exact.astype(np.complex128) - born) ** 2) /
                np.sum(np.abs(born.astype(np.complex128) - 1) ** 2))
    assert e < 0.02


def test_inversion_of_exact_bead_known_answer(oracle, qh):
    """The reference's measured values (tests/test_bpm.cpp:166-167): RMSE 6.04e-3 (Rytov + GP)
    -> 3.21e-3, inner mean 1.4030 -> 1.4019, 20 losses, the last below a tenth of the first."""
    _, il, data, truth, initial, inner = exact_bead_case(oracle, qh)
    ri, loss = oracle.bpm_reconstruct(data, il, GEOM, initial)

    def rmse(v):
        return float(np.sqrt(np.mean((v.astype(np.float64) - truth) ** 2)))

    assert rmse(initial) == pytest.approx(6.04e-3, abs=5e-5)
    assert rmse(ri) == pytest.approx(3.21e-3, abs=5e-5)
    assert rmse(ri) < 0.6 * rmse(initial)
    assert float(initial[inner].mean()) == pytest.approx(1.4030, abs=1e-4)
    assert float(ri[inner].mean()) == pytest.approx(1.4019, abs=1e-4)
    assert loss.shape == (20,) and loss[-1] < 0.1 * loss[0]
