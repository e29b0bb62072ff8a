"""Beam propagation on the GPU against the oracle: forward model, adjoint (vjp) with slice
checkpointing, gradient and loss, the inversion, the torch.autograd.Function, and invariants
at the full 256^2 x 256 / 49-illumination shape (unit envelope, checkpoint and chunk
invariance bit for bit, linearity of the adjoint)."""

import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

GEOM = (0.532, 1.337, 1.2, 0.2)


def spheres(ph):
    """The two nested spheres of the reference's tests/test_cuda.cpp:30-31."""
    return [ph.Sphere(0.0, 0.4, -0.3, 2.5, 1.36), ph.Sphere(0.0, 0.4, -0.3, 1.0, 1.38)]


def residual_pattern(count, n):
    i = np.arange(count * n * n, dtype=np.float64)
    return (np.sin(0.7 * i) + 1j * np.cos(1.3 * i)).astype(np.complex64).reshape(count, n, n)


def case(qh, nz, n, count=7, na=0.9):
    from qoraholo_amd import phantom as ph

    g = qh.Geometry(*GEOM)
    il = qh.circular_scan(count, na, g)
    ri = ph.render_phantom(spheres(ph), g, nz, n)
    return g, il, ri, (ri - np.float32(GEOM[1])).astype(np.float32)


@pytest.mark.parametrize("nz,n", [(32, 48), (16, 32), (24, 64), (12, 128), (6, 256), (4, 512),
                                  (10, 60), (8, 192), (12, 96), (4, 384), (6, 80)])
def test_forward_matches_oracle(device, qh, oracle, nz, n):
    """Scattered part y - 1 against the FP64 oracle: as close as the float oracle (the
    reference's float algorithm) is, within a factor 2. At large n with a small object the
    scattered part is small against 1 and float rounding of y sets the floor for both."""
    g, il, ri, x = case(qh, nz, n)
    print(f"bpm forward nz={nz} n={n}", flush=True)
    for obliquity in (True, False):
        got = qh.bpm_forward(ri, il, g, obliquity=obliquity)
        f32 = oracle.bpm_forward(x, il, GEOM, obliquity=obliquity)
        f64 = oracle.bpm_forward(x, il, GEOM, obliquity=obliquity, dtype=np.float64)
        assert got.shape == f64.shape
        e_ref = rel_l2(f32 - 1, f64 - 1)
        e_gpu = rel_l2(got - 1, f64 - 1)
        assert e_gpu < max(2 * e_ref, 1e-5), (obliquity, e_gpu, e_ref)
        assert rel_l2(got - 1, f32 - 1) < 5e-5


def test_forward_chunks_are_bitwise_equal(device, qh):
    g, il, ri, _ = case(qh, 32, 48)
    a = qh.bpm_forward(ri, il, g)
    for chunk in (1, 3):
        assert np.array_equal(qh.bpm_forward(ri, il, g, chunk=chunk), a)


@pytest.mark.parametrize("nz,n", [(32, 48), (16, 32), (20, 64)])
def test_vjp_matches_oracle(device, qh, oracle, nz, n):
    g, il, _, x = case(qh, nz, n)
    r = residual_pattern(il.shape[0], n)
    print(f"bpm vjp nz={nz} n={n}", flush=True)
    got = qh.bpm_vjp(x, r, il, g)
    want = oracle.bpm_vjp(x, r, il, GEOM)
    # The reference's CUDA-vs-CPU bound is 1e-4 (tests/test_cuda.cpp:149-160).
    assert rel_l2(got, want) < 2e-5


def test_vjp_invariant_to_checkpoints_and_chunks(device, qh):
    """Recomputed envelopes are the stored ones bit for bit, and the gradient terms are summed
    in illumination order whatever the chunking."""
    g, il, _, x = case(qh, 32, 48)
    r = residual_pattern(il.shape[0], 48)
    ref = qh.bpm_vjp(x, r, il, g, checkpoint=0)
    for checkpoint, chunk in ((16, 0), (5, 0), (1, 0), (0, 2), (7, 3), (-1, 0), (-1, 3)):
        assert np.array_equal(qh.bpm_vjp(x, r, il, g, checkpoint=checkpoint, chunk=chunk), ref), (
            checkpoint, chunk)
    with pytest.raises(ValueError, match="checkpoint at least -1"):
        qh.bpm_vjp(x, r, il, g, checkpoint=-2)


@pytest.mark.parametrize("nz,n,count", [(16, 32, 5), (24, 48, 7), (12, 64, 4), (8, 96, 3),
                                         (6, 256, 2)])
def test_jvp_matches_oracle(device, qh, oracle, nz, n, count):
    """Device tangent model (bpm_jvp, src/cpu/bpm.cpp:141-176) against the float oracle's and
    against the FP64 one, where the float oracle sets the tolerance: as close as the float
    oracle is, within a factor 2."""
    g, il, _, x = case(qh, nz, n, count=count)
    rng = np.random.default_rng(nz * n)
    h = rng.standard_normal(x.shape).astype(np.float32)
    print(f"bpm jvp nz={nz} n={n} count={count}", flush=True)
    got = qh.bpm_jvp(x, h, il, g)
    f32 = oracle.bpm_jvp(x, h, il, GEOM)
    f64 = oracle.bpm_jvp(x, h, il, GEOM, dtype=np.float64)
    assert got.shape == f64.shape
    assert rel_l2(got, f32) < 2e-5
    assert rel_l2(got, f64) < max(2 * rel_l2(f32, f64), 1e-5)
    # Chunked the same bits.
    assert np.array_equal(qh.bpm_jvp(x, h, il, g, chunk=1), got)


def test_jvp_adjoint_identity_and_difference_on_device(device, qh):
    """<J h, r> = <h, J^T r> with both sides from the device, and J h against a central
    difference of the device forward model (the checks of the reference's
    tests/test_bpm.cpp:88-126); the difference at eps 1e-3 is float-limited, the reference
    documents 1.75e-4 for its CPU model."""
    g, il, _, x = case(qh, 16, 32, count=5)
    rng = np.random.default_rng(5)
    h = rng.standard_normal(x.shape).astype(np.float32)
    r = residual_pattern(il.shape[0], 32)
    jh = qh.bpm_jvp(x, h, il, g).astype(np.complex128)
    lhs = np.sum(np.real(np.conj(jh) * r.astype(np.complex128)))
    rhs = np.sum(h.astype(np.float64) * qh.bpm_vjp(x, r, il, g).astype(np.float64))
    assert abs(lhs - rhs) / abs(rhs) < 2e-5
    eps = np.float32(1e-3)
    from qoraholo_amd.bpm import bpm_forward_contrast

    fp = bpm_forward_contrast(x + eps * h, il, g).astype(np.complex128)
    fm = bpm_forward_contrast(x - eps * h, il, g).astype(np.complex128)
    fd = (fp - fm) / (2 * 1e-3)
    assert rel_l2(fd, jh) < 2e-3


def test_table_entry_points_equal_geometry_entry_points(device, qh):
    """The calls on BpmTables (the reference's backend signature, src/backends.hpp:119-124)
    give the same bits as the calls that build the tables from the illuminations."""
    g, il, ri, x = case(qh, 16, 48, count=5)
    t = qh.BpmTables.build(48, 16, il, g)
    r = residual_pattern(il.shape[0], 48)
    rng = np.random.default_rng(9)
    h = rng.standard_normal(x.shape).astype(np.float32)
    assert np.array_equal(t.forward(x), qh.bpm_forward(ri, il, g))
    assert np.array_equal(t.vjp(x, r), qh.bpm_vjp(x, r, il, g))
    assert np.array_equal(t.jvp(x, h), qh.bpm_jvp(x, h, il, g))
    fields = qh.bpm_forward(ri, il, g)
    start = np.full_like(x, np.float32(0.0))
    got, loss = t.reconstruct(fields, start, iterations=3, tv_weight=1e-2)
    want, want_loss = qh.bpm_reconstruct(fields, il, g, start + np.float32(GEOM[1]), iterations=3,
                                         tv_weight=1e-2, return_loss=True)
    assert np.array_equal(got + np.float32(GEOM[1]), want)
    assert np.array_equal(loss, want_loss)


def test_jvp_dev_equals_host_call(device, qh):
    import torch

    g, il, _, x = case(qh, 16, 32, count=5)
    rng = np.random.default_rng(6)
    h = rng.standard_normal(x.shape).astype(np.float32)
    plan = qh.BpmPlan(16, 32, il, g)
    got = plan.jvp(torch.from_numpy(x).cuda(), torch.from_numpy(h).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(got.cpu().numpy(), qh.bpm_jvp(x, h, il, g))


def test_adjoint_identity_on_device(device, qh, oracle):
    """<J h, r> with the FP64 tangent model against <h, J^T r> from the device adjoint."""
    g, il, _, x = case(qh, 16, 32, count=5)
    rng = np.random.default_rng(3)
    h = rng.standard_normal(x.shape).astype(np.float32)
    r = residual_pattern(il.shape[0], 32)
    jh = oracle.bpm_jvp(x, h, il, GEOM, dtype=np.float64)
    lhs = np.sum(np.real(np.conj(jh) * r.astype(np.complex128)))
    rhs = np.sum(h.astype(np.float64) * qh.bpm_vjp(x, r, il, g).astype(np.float64))
    assert abs(lhs - rhs) / abs(rhs) < 2e-5


def test_gradient_against_fp64_differences(device, qh, oracle):
    """Central differences (step 1e-4) of the FP64 oracle's loss at 32 seeded voxels against
    the device gradient."""
    torch = pytest.importorskip("torch")
    g, il, _, x = case(qh, 16, 32, count=5)
    data = oracle.bpm_forward((x * 1.2).astype(np.float32), il, GEOM)
    plan = qh.BpmPlan(16, 32, il, g)
    xt = torch.from_numpy(x).cuda()
    dt = torch.from_numpy(data).cuda()
    grad, loss = plan.gradient(xt, dt)
    grad = grad.cpu().numpy()
    x64 = x.astype(np.float64)
    _, l64 = oracle.bpm_gradient(x64, data, il, GEOM, dtype=np.float64)
    assert loss == pytest.approx(l64, rel=1e-5)
    rng = np.random.default_rng(5)
    eps = 1e-4
    for i in rng.choice(x.size, 32, replace=False):
        xp = x64.copy().reshape(-1)
        xm = x64.copy().reshape(-1)
        xp[i] += eps
        xm[i] -= eps
        _, lp = oracle.bpm_gradient(xp.reshape(x.shape), data, il, GEOM, dtype=np.float64)
        _, lm = oracle.bpm_gradient(xm.reshape(x.shape), data, il, GEOM, dtype=np.float64)
        fd = (lp - lm) / (2 * eps)
        assert grad.reshape(-1)[i] == pytest.approx(fd, rel=2e-3, abs=1e-4 * np.abs(grad).max())


def test_plan_gradient_matches_oracle(device, qh, oracle):
    torch = pytest.importorskip("torch")
    g, il, _, x = case(qh, 32, 48)
    data = oracle.bpm_forward((x * 0.8).astype(np.float32), il, GEOM)
    plan = qh.BpmPlan(32, 48, il, g)
    xt = torch.from_numpy(x).cuda()
    dt = torch.from_numpy(data).cuda()
    for first, last in ((0, 7), (2, 5)):
        grad, loss = plan.gradient(xt, dt, first, last)
        want, wl = oracle.bpm_gradient(x, data, il, GEOM, first=first, last=last)
        assert rel_l2(grad.cpu().numpy(), want) < 2e-5
        assert loss == pytest.approx(wl, rel=2e-5)


def test_autograd_function(device, qh):
    """torch.autograd through the plan: the backward of 0.5 sum |bpm(x) - d|^2 is the adjoint
    of the residual, count times the plan's gradient (which carries 1 / count)."""
    torch = pytest.importorskip("torch")
    g, il, _, x = case(qh, 16, 64)
    plan = qh.BpmPlan(16, 64, il, g)
    xt = torch.from_numpy(x).cuda()
    d = plan.forward(xt * 0.9)
    xr = xt.clone().requires_grad_(True)
    y = qh.bpm(xr, plan)
    loss = 0.5 * torch.sum(torch.abs(y - d) ** 2)
    loss.backward()
    grad, l2 = plan.gradient(xt, d)
    count = il.shape[0]
    assert rel_l2(xr.grad.cpu().numpy(), count * grad.cpu().numpy()) < 1e-5
    assert loss.item() / count == pytest.approx(l2, rel=1e-5)
    direct = plan.vjp(xt, (y - d).detach().contiguous())
    assert rel_l2(direct.cpu().numpy(), xr.grad.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("batch,tv", [(0, 0.0), (3, 0.0), (0, 1e-2), (3, 1e-2)])
def test_reconstruct_matches_oracle(device, qh, oracle, batch, tv):
    """tests/test_cuda.cpp:163-183 (5 iterations from a flat start, with and without batches and
    TV); the reference's bound is 1e-3 on the contrast and on the last loss."""
    g, il, ri, x = case(qh, 32, 48)
    data = oracle.bpm_forward(x, il, GEOM)
    start = np.full_like(ri, np.float32(GEOM[1]))
    print(f"bpm reconstruct batch={batch} tv={tv}", flush=True)
    got, gl = qh.bpm_reconstruct(data, il, g, start, iterations=5, batch=batch, tv_weight=tv,
                                 return_loss=True)
    want, wl = oracle.bpm_reconstruct(data, il, GEOM, start, iterations=5, batch=batch,
                                      tv_weight=tv)
    nm = np.float32(GEOM[1])
    assert rel_l2(got - nm, want - nm) < 1e-4
    assert np.allclose(gl, wl, rtol=1e-4)
    assert gl[-1] < gl[0]


def test_reconstruct_chunk_invariant(device, qh, oracle):
    """The gradient computed in device passes of 2 illuminations gives the same result
    (tests/test_cuda.cpp:186-196; here bit for bit)."""
    g, il, ri, x = case(qh, 32, 48)
    data = oracle.bpm_forward(x, il, GEOM)
    start = np.full_like(ri, np.float32(GEOM[1]))
    a = qh.bpm_reconstruct(data, il, g, start, iterations=3)
    b = qh.bpm_reconstruct(data, il, g, start, iterations=3, chunk=2, checkpoint=8)
    assert np.array_equal(a, b)


def test_full_size_invariants(device, qh):
    """256^2 x 256 with 49 illuminations: the unit envelope passes through x = 0 exactly,
    checkpointed and stored envelopes give the same adjoint bit for bit, the automatic
    interval keeps every envelope (6.6 GB, far below half of the free memory), and the adjoint
    is linear in r."""
    torch = pytest.importorskip("torch")
    g = qh.Geometry(*GEOM)
    il = qh.circular_scan(49, 0.9, g)
    n = nz = 256
    print("bpm full size 256^3 x 49", flush=True)
    plan = qh.BpmPlan(nz, n, il, g, checkpoint=16)
    assert plan.checkpoint == 16
    zero = torch.zeros((nz, n, n), dtype=torch.float32, device="cuda")
    y0 = plan.forward(zero)
    assert torch.equal(y0, torch.ones_like(y0))
    gen = torch.Generator(device="cuda").manual_seed(7)
    x = 0.02 * torch.rand((nz, n, n), generator=gen, device="cuda")
    r1 = torch.randn((49, n, n), dtype=torch.complex64, generator=gen, device="cuda")
    r2 = torch.randn((49, n, n), dtype=torch.complex64, generator=gen, device="cuda")
    g1 = plan.vjp(x, r1)
    g2 = plan.vjp(x, r2)
    g12 = plan.vjp(x, (r1 + r2).contiguous())
    lin = float(torch.linalg.vector_norm(g12 - g1 - g2) / torch.linalg.vector_norm(g12))
    assert lin < 1e-5
    stored = qh.BpmPlan(nz, n, il, g, checkpoint=0)
    assert stored.checkpoint == 0
    assert torch.equal(stored.vjp(x, r1), g1)
    del stored
    auto = qh.BpmPlan(nz, n, il, g)
    assert auto.checkpoint == 0 and auto.chunk == 49
    assert torch.equal(auto.vjp(x, r1), g1)
    del auto
    torch.cuda.synchronize()


def test_automatic_checkpoint_with_little_free_memory(device, qh):
    """256^2 x 256 with 49 illuminations while all but 8 GiB of the device memory is taken: the
    envelopes (6.6 GB) would not fit within half the free memory, so the automatic interval
    falls back to checkpoints every 16 slices, still in one pass of all 49 illuminations, and the
    adjoint has the bits of a plan that keeps every envelope."""
    torch = pytest.importorskip("torch")
    g = qh.Geometry(*GEOM)
    il = qh.circular_scan(49, 0.9, g)
    n = nz = 256
    gen = torch.Generator(device="cuda").manual_seed(3)
    x = 0.02 * torch.rand((nz, n, n), generator=gen, device="cuda")
    r = torch.randn((49, n, n), dtype=torch.complex64, generator=gen, device="cuda")
    torch.cuda.synchronize()
    free, _ = torch.cuda.mem_get_info()
    print(f"bpm automatic checkpoint: {free / 2**30:.1f} GiB free, keeping 8", flush=True)
    assert free > 16 << 30
    hog = torch.empty(free - (8 << 30), dtype=torch.uint8, device="cuda")
    plan = qh.BpmPlan(nz, n, il, g)
    assert plan.checkpoint == 16 and plan.chunk == 49
    g16 = plan.vjp(x, r)
    torch.cuda.synchronize()
    del plan, hog
    torch.cuda.empty_cache()
    stored = qh.BpmPlan(nz, n, il, g, checkpoint=0)
    assert torch.equal(stored.vjp(x, r), g16)
    del stored
    torch.cuda.synchronize()


def overlapping_beads(ph):
    """Three overlapping beads, radii 2-4 um, Delta n up to 0.048 above the medium."""
    return [ph.Sphere(0.0, 0.0, 0.0, 4.0, 1.385), ph.Sphere(2.5, 1.5, -1.0, 3.0, 1.372),
            ph.Sphere(-2.0, -2.0, 2.0, 2.0, 1.360)]


def test_gradient_at_full_size(device, qh, oracle):
    """The BPM adjoint workload at its own size: 256^2 x 256 at 0.2 um, 49 illuminations at NA
    0.9, three overlapping beads, checkpoints every 16 slices. The data-term gradient at 0.7
    times the true contrast, on a subset of illuminations [first, last) (the device plan holds
    all 49), against the oracle's float gradient on two subsets, and against central
    differences (step 1e-4) of the FP64 oracle's loss on illumination 0 at 64 seeded voxels
    (the reference's pattern, tests/test_cuda.cpp:149-160 and tests/test_bpm.cpp:108-126)."""
    torch = pytest.importorskip("torch")
    from concurrent.futures import ThreadPoolExecutor

    from qoraholo_amd import phantom as ph

    n = nz = 256
    g = qh.Geometry(*GEOM)
    il = qh.circular_scan(49, 0.9, g)
    x_true = (ph.render_phantom(overlapping_beads(ph), g, nz, n) - np.float32(GEOM[1]))
    x_true = x_true.astype(np.float32)
    print("bpm gradient at full size 256^3 x 49", flush=True)
    plan = qh.BpmPlan(nz, n, il, g, checkpoint=16)
    data = plan.forward(torch.from_numpy(x_true).cuda()).cpu().numpy()
    x = (0.7 * x_true).astype(np.float32)
    xt = torch.from_numpy(x).cuda()
    dt = torch.from_numpy(data).cuda()
    for first, last in ((0, 2), (30, 33)):
        grad, _ = plan.gradient(xt, dt, first, last)
        want, _ = oracle.bpm_gradient(x, data, il, GEOM, first=first, last=last)
        err = rel_l2(grad.cpu().numpy(), want)
        print(f"  illuminations [{first}, {last}): rel L2 vs oracle {err:.2e}", flush=True)
        assert err < 2e-5
    grad, _ = plan.gradient(xt, dt, 0, 1)
    grad = grad.cpu().numpy().reshape(-1)
    del plan
    torch.cuda.synchronize()

    x64 = x.astype(np.float64)
    d0 = data[:1].astype(np.complex128)
    il0 = il[:1]

    def loss(i, sign):
        oracle.set_threads(1)
        xp = x64.copy().reshape(-1)
        xp[i] += sign * 1e-4
        y = oracle.bpm_forward(xp.reshape(x.shape), il0, GEOM, dtype=np.float64)
        return 0.5 * float(np.sum(np.abs(y - d0) ** 2))

    voxels = np.random.default_rng(11).choice(x.size, 64, replace=False)
    jobs = [(int(i), s) for i in voxels for s in (1.0, -1.0)]
    with ThreadPoolExecutor(max_workers=min(16, len(os.sched_getaffinity(0)))) as ex:
        values = list(ex.map(lambda job: loss(*job), jobs))
    scale = np.abs(grad).max()
    for k, i in enumerate(voxels):
        fd = (values[2 * k] - values[2 * k + 1]) / 2e-4
        assert grad[i] == pytest.approx(fd, rel=2e-3, abs=1e-4 * scale), (int(i), grad[i], fd)


def test_served_request_at_384(device, qh, oracle):
    """The served BPM inversion at its larger size: 384^2 x 384 with 81 illuminations through
    the Server (Rytov, GP 20 start, 2 FISTA iterations with the TV prox of weight 1e-2). The
    served result equals the per-call chain bit for bit; the loss falls and n >= n_m; and the
    device gradient at the start, on 2 of the 81 illuminations, matches the oracle's, with every
    envelope stored (the automatic interval at this size) and, bit for bit, with checkpoints
    every 16 slices."""
    torch = pytest.importorskip("torch")
    from qoraholo_amd import workloads

    req = workloads.bpm_request(384, 81, seed=0)
    g, il, f = req.geometry, req.illuminations, req.fields
    nm = np.float32(GEOM[1])
    print("bpm served request 384^3 x 81", flush=True)
    server = qh.Server(max_slots=1)
    gid = server.add_bpm(384, il, g, start_iterations=20, iterations=2, tv_weight=1e-2,
                         tv_inner=10)
    served, served_loss, _ = server.wait(server.submit_bpm(gid, f))
    del server
    start = qh.reconstruct(qh.rytov(f), il, g, iterations=20)
    ri, loss = qh.bpm_reconstruct(f, il, g, start, iterations=2, tv_weight=1e-2, tv_inner=10,
                                  return_loss=True)
    assert np.array_equal(served, ri)
    assert np.array_equal(served_loss, loss)
    assert loss[-1] < loss[0]
    assert float(ri.min()) >= nm

    x = (start - nm).astype(np.float32)
    plan = qh.BpmPlan(384, 384, il, g)
    assert plan.checkpoint == 0 and plan.chunk == 81  # 36.7 GB of envelopes fit
    xt, ft = torch.from_numpy(x).cuda(), torch.from_numpy(f).cuda()
    grad, _ = plan.gradient(xt, ft, 40, 42)
    del plan
    plan = qh.BpmPlan(384, 384, il, g, checkpoint=16)
    assert torch.equal(plan.gradient(xt, ft, 40, 42)[0], grad)  # 24 recomputed segments
    del plan
    want, _ = oracle.bpm_gradient(x, f, il, GEOM, first=40, last=42)
    err = rel_l2(grad.cpu().numpy(), want)
    print(f"  gradient on illuminations [40, 42): rel L2 vs oracle {err:.2e}", flush=True)
    assert err < 2e-5


def test_inversion_of_exact_bead_on_device(device, qh, oracle):
    """The reference's exact-bead inversion (tests/test_bpm.cpp:131-171) with the device path
    end to end (Rytov, GP 20 start, 20 BPM iterations): the reference's measured values and the
    oracle's volume."""
    from test_oracle_bpm import exact_bead_case

    g, il, data, truth, initial_o, inner = exact_bead_case(oracle, qh)
    initial = qh.reconstruct(qh.rytov(data), il, g, iterations=20)
    nm = np.float32(GEOM[1])
    assert rel_l2(initial - nm, initial_o - nm) < 5e-5
    ri, loss = qh.bpm_reconstruct(data, il, g, initial, return_loss=True)
    want, wl = oracle.bpm_reconstruct(data, il, GEOM, initial_o)
    assert rel_l2(ri - nm, want - nm) < 1e-4
    assert np.allclose(loss, wl, rtol=1e-4)
    rmse = float(np.sqrt(np.mean((ri.astype(np.float64) - truth) ** 2)))
    assert rmse == pytest.approx(3.21e-3, abs=5e-5)
    assert float(ri[inner].mean()) == pytest.approx(1.4019, abs=1e-4)


def test_served_request_at_192_matches_oracle(device, qh, oracle):
    """One inversion request of the served configuration at 192^2 x 192 with 49 illuminations
    (a random cell at 0.2 um): Rytov + GP 20 start on the device, then 2 FISTA iterations with
    the TV prox (weight 1e-2), against the oracle from the same start and data."""
    from qoraholo_amd import workloads

    req = workloads.bpm_request(192, 49, seed=0)
    g, il, f = req.geometry, req.illuminations, req.fields
    assert (g.wavelength, g.n_medium, g.na, g.pixel) == GEOM
    print("bpm served request 192^3 x 49", flush=True)
    start = qh.reconstruct(qh.rytov(f), il, g, iterations=20)
    got, gl = qh.bpm_reconstruct(f, il, g, start, iterations=2, tv_weight=1e-2,
                                 return_loss=True)
    want, wl = oracle.bpm_reconstruct(f, il, GEOM, start, iterations=2, tv_weight=1e-2)
    nm = np.float32(GEOM[1])
    assert rel_l2(got - nm, want - nm) < 1e-4
    assert np.allclose(gl, wl, rtol=1e-4)
    assert gl[-1] < gl[0]
    assert float(got.min()) >= nm


@pytest.mark.parametrize("case", ["bead", "stack"])
def test_bpm_figure_reproduces_published_tables(device, qh, case):
    """The reference README's BPM tables (README.md:182-204, examples/bpm_figure.py): 128^3 at
    0.15 um, 49 illuminations, no noise; bead data from the exact partial-wave model, stack data
    from the BPM on a grid twice as fine cut to the coarse grid in Fourier space. Rytov, GP 20,
    FISTA-TV 20x10, and 30 BPM iterations from Rytov + GP (plain, TV 3e-2, plain phase screen),
    all on the device: every published RMSE and inner mean to the printed digits."""
    published = {
        "bead": [(6.87e-3, 1.3882), (5.36e-3, 1.4016), (4.93e-3, 1.4009), (2.50e-3, 1.4019),
                 (2.39e-3, 1.4018), (5.06e-3, 1.4279)],
        "stack": [(4.94e-3, 1.3807), (4.27e-3, 1.3964), (3.89e-3, 1.3990), (2.00e-3, 1.4031),
                  (1.95e-3, 1.4032), (3.64e-3, 1.4268)],
    }[case]
    n, angles = 128, 49
    geom = qh.Geometry(0.532, 1.337, 1.2, 0.15)
    illum = qh.circular_scan(angles, 0.9, geom)
    print(f"bpm figure {case}", flush=True)
    if case == "bead":
        spheres = [qh.Sphere(0.0, 0.0, 0.0, 4.0, 1.40)]
        data = qh.forward_fields(spheres, illum, geom, n, model="exact")
    else:
        spheres = [qh.Sphere(z, 0.0, 0.0, 2.0, 1.40) for z in (-4.5, 0.0, 4.5)]
        fine = qh.Geometry(0.532, 1.337, 1.2, 0.075)
        f = qh.bpm_forward(qh.render_phantom(spheres, fine, 2 * n, 2 * n), illum, fine)
        spec = np.fft.fftshift(np.fft.fft2(f), axes=(1, 2))
        lo = n // 2
        crop = np.fft.ifftshift(spec[:, lo:lo + n, lo:lo + n], axes=(1, 2))
        data = (np.fft.ifft2(crop) / 4).astype(np.complex64)
    truth = qh.render_phantom(spheres, geom, n, n)
    z, y, x = (np.indices(truth.shape) - n // 2) * geom.pixel
    inner = np.zeros(truth.shape, bool)
    for s in spheres:
        inner |= (z - s.z) ** 2 + (y - s.y) ** 2 + (x - s.x) ** 2 < (0.6 * s.radius) ** 2
    ry = qh.rytov(data)
    gp = qh.reconstruct(ry, illum, geom, iterations=20)
    vols = [qh.reconstruct(ry, illum, geom), gp,
            qh.reconstruct(ry, illum, geom, iterations=20, solver="tv"),
            qh.bpm_reconstruct(data, illum, geom, gp, iterations=30),
            qh.bpm_reconstruct(data, illum, geom, gp, iterations=30, tv_weight=3e-2),
            qh.bpm_reconstruct(data, illum, geom, gp, iterations=30, obliquity=False)]
    for v, (rmse, mean) in zip(vols, published):
        got = float(np.sqrt(np.mean((v - truth) ** 2)))
        assert float(f"{got:.2e}") == rmse, (got, rmse)
        assert round(float(v[inner].mean()), 4) == mean, (float(v[inner].mean()), mean)
