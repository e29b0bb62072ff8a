"""Served loop: every request's result is the one it gets alone, whatever else is in flight;
admission against the HBM budget; error isolation; latency statistics."""

import threading

import numpy as np
import pytest

from test_gpu_parity import PipelineData

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mixed(qh):
    """Three geometries of different sizes and solvers, with one alone result each."""
    cases = []
    for holo, angles, iterations, solver in ((128, 9, 5, "gp"), (128, 9, 4, "tv"),
                                             (256, 25, 3, "tv")):
        d = PipelineData(qh, holo=holo, angles=angles, iterations=iterations, solver=solver)
        p = qh.Pipeline(d.count, d.size, illuminations=d.illum, geometry=d.geom, **d.kw)
        p.set_background(d.bg)
        alone = p.run(d.holo)
        cases.append((d, alone))
    return cases


def add_all(server, cases):
    return [server.add_geometry(d.count, d.size, illuminations=d.illum, geometry=d.geom,
                                backgrounds=d.bg, **d.kw) for d, _ in cases]


@pytest.mark.parametrize("max_slots", [4, 1])
def test_results_are_the_same_alone_and_in_the_mix(device, qh, mixed, max_slots):
    server = qh.Server(max_slots=max_slots)
    gids = add_all(server, mixed)
    order = [0, 2, 1, 0, 1, 2, 2, 0, 1]
    ids = [(k, server.submit(gids[k], mixed[k][0].holo)) for k in order]
    for k, rid in ids:
        ri, filled, latency = server.wait(rid)
        assert np.array_equal(ri, mixed[k][1]), (k, max_slots)
        assert latency > 0
    s = server.stats()
    assert s["completed"] == len(order) and s["failed"] == 0
    assert 0 < s["p50_ms"] <= s["p99_ms"]
    assert 1 <= s["peak_busy_slots"] <= max_slots
    # Idle slots stay resident: each of the three geometries keeps one (the budget is ample).
    assert s["slots"] >= len(mixed)
    # Every reconstruction slot captured its graph when it was built and replays it.
    assert s["graph_slots"] == s["slots"]


def test_submissions_from_several_threads(device, qh, mixed):
    server = qh.Server(max_slots=3)
    gids = add_all(server, mixed)
    results = {}

    def client(k):
        rid = server.submit(gids[k], mixed[k][0].holo)
        results[(k, rid)] = server.wait(rid)[0]

    threads = [threading.Thread(target=client, args=(k,)) for k in (0, 1, 2, 0, 2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert len(results) == 5
    for (k, _), ri in results.items():
        assert np.array_equal(ri, mixed[k][1])


def test_admission_refuses_a_geometry_over_budget(device, qh, mixed):
    server = qh.Server(budget_bytes=1 << 20)
    d = mixed[0][0]
    with pytest.raises(qh.OutOfDeviceMemory):
        server.add_geometry(d.count, d.size, illuminations=d.illum, geometry=d.geom,
                            backgrounds=d.bg, **d.kw)


def test_bad_requests_are_rejected_alone(device, qh, mixed):
    server = qh.Server(max_slots=2)
    gids = add_all(server, mixed[:1])
    d, alone = mixed[0]
    with pytest.raises(ValueError):
        server.submit(7, d.holo)
    with pytest.raises(ValueError):
        server.submit(gids[0], d.holo[:, :64])
    rid = server.submit(gids[0], d.holo)
    assert np.array_equal(server.wait(rid)[0], alone)


# ---- BPM inversion requests (the served BPM configuration at test sizes) ----

BPM_CASES = ((48, 9, 3, 0.0), (64, 13, 4, 1e-2))  # n, illuminations, iterations, TV weight


@pytest.fixture(scope="module")
def bpm_mixed(qh):
    """Two BPM geometries and, for each, the result of the per-call chain the served request
    replaces (examples/bpm_figure.py:74-88): rytov, a 2-iteration GP start, bpm_reconstruct."""
    from qoraholo_amd import workloads

    cases = []
    for n, angles, iterations, tv in BPM_CASES:
        req = workloads.bpm_request(n, angles, seed=n)
        g, il, f = req.geometry, req.illuminations, req.fields
        start = qh.reconstruct(qh.rytov(f), il, g, iterations=2)
        ri, loss = qh.bpm_reconstruct(f, il, g, start, iterations=iterations, tv_weight=tv,
                                      tv_inner=4, return_loss=True)
        kw = dict(start_iterations=2, iterations=iterations, tv_weight=tv, tv_inner=4)
        cases.append((req, kw, ri, loss))
    return cases


def test_bpm_requests_equal_the_per_call_chain_alone_and_in_the_mix(device, qh, mixed, bpm_mixed):
    """Each served BPM inversion gives the per-call chain's index volume and loss bit for bit,
    alone and interleaved with another BPM geometry and reconstruction requests."""
    server = qh.Server(max_slots=4)
    bids = [server.add_bpm(req.fields.shape[1], req.illuminations, req.geometry, **kw)
            for req, kw, _, _ in bpm_mixed]
    print("served bpm alone", flush=True)
    for k, gid in enumerate(bids):
        ri, loss, latency = server.wait(server.submit_bpm(gid, bpm_mixed[k][0].fields))
        assert np.array_equal(ri, bpm_mixed[k][2]), k
        assert np.array_equal(loss, bpm_mixed[k][3]), k
        assert latency > 0
    print("served bpm mixed", flush=True)
    rids = add_all(server, mixed[:2])
    order = [("b", 1), ("r", 0), ("b", 0), ("b", 1), ("r", 1), ("b", 0)]
    ids = []
    for kind, k in order:
        if kind == "b":
            ids.append((kind, k, server.submit_bpm(bids[k], bpm_mixed[k][0].fields)))
        else:
            ids.append((kind, k, server.submit(rids[k], mixed[k][0].holo)))
    for kind, k, rid in ids:
        out, extra, _ = server.wait(rid)
        if kind == "b":
            assert np.array_equal(out, bpm_mixed[k][2]), (kind, k)
            assert np.array_equal(extra, bpm_mixed[k][3]), (kind, k)
        else:
            assert np.array_equal(out, mixed[k][1]), (kind, k)
    s = server.stats()
    assert s["completed"] == len(order) + len(bids) and s["failed"] == 0
    assert s["peak_busy_slots"] <= 4


def test_bpm_requests_are_checked(device, qh, bpm_mixed):
    server = qh.Server(max_slots=2)
    req, kw, ri, _ = bpm_mixed[0]
    gid = server.add_bpm(req.fields.shape[1], req.illuminations, req.geometry, **kw)
    with pytest.raises(ValueError):
        server.submit_bpm(gid, req.fields[:, :16])
    with pytest.raises(ValueError):
        server.submit(gid, req.fields.real)
    with pytest.raises(ValueError):
        server.add_bpm(req.fields.shape[1], req.illuminations, req.geometry, iterations=2,
                       step=-1.0)
    rid = server.submit_bpm(gid, req.fields)
    assert np.array_equal(server.wait(rid)[0], ri)


def test_bpm_geometry_over_budget_is_refused(device, qh, bpm_mixed):
    server = qh.Server(budget_bytes=1 << 20)
    req, kw, _, _ = bpm_mixed[0]
    with pytest.raises(qh.OutOfDeviceMemory):
        server.add_bpm(req.fields.shape[1], req.illuminations, req.geometry, **kw)


def test_c_abi_refuses_a_request_of_the_other_kind(device, qh, mixed, bpm_mixed):
    """The C entry points check the geometry's request kind (the Python layer checks first)."""
    import ctypes

    from qoraholo_amd import _native

    server = qh.Server(max_slots=2)
    req, kw, _, _ = bpm_mixed[0]
    bpm = server.add_bpm(req.fields.shape[1], req.illuminations, req.geometry, **kw)
    rec = add_all(server, mixed[:1])[0]
    out = np.empty((64, 64, 64), np.float32)
    rid = ctypes.c_int64()
    holo = mixed[0][0].holo
    with pytest.raises(ValueError):
        _native.check(_native.lib().qh_hip_server_submit(server._handle, bpm, _native.ptr(holo),
                                                         _native.ptr(out), ctypes.byref(rid)))
    with pytest.raises(ValueError):
        _native.check(_native.lib().qh_hip_server_submit_bpm(
            server._handle, rec, _native.ptr(req.fields), _native.ptr(out), None,
            ctypes.byref(rid)))
    assert server.stats()["submitted"] == 0


def test_tight_budget_evicts_across_request_kinds(device, qh, mixed, bpm_mixed):
    """With room for one slot at a time, a reconstruction geometry and a BPM geometry take turns:
    registering the second evicts the first's idle slot, and every request evicts the other
    kind's; each result is still the one the request gets alone."""
    probe = qh.Server(max_slots=4)
    req, kw, ri_bpm, loss_bpm = bpm_mixed[0]
    probe.add_bpm(req.fields.shape[1], req.illuminations, req.geometry, **kw)
    bpm_bytes = probe.stats()["device_bytes"]
    add_all(probe, mixed[:1])
    rec_bytes = probe.stats()["device_bytes"] - bpm_bytes
    del probe
    budget = max(bpm_bytes, rec_bytes) + min(bpm_bytes, rec_bytes) // 2  # one slot at a time
    server = qh.Server(budget_bytes=budget, max_slots=4)
    bpm = server.add_bpm(req.fields.shape[1], req.illuminations, req.geometry, **kw)
    rec = add_all(server, mixed[:1])[0]
    assert server.stats()["slots"] == 1
    for kind in ("r", "b", "r", "b", "b", "r"):
        if kind == "b":
            ri, loss, _ = server.wait(server.submit_bpm(bpm, req.fields))
            assert np.array_equal(ri, ri_bpm) and np.array_equal(loss, loss_bpm)
        else:
            ri, _, _ = server.wait(server.submit(rec, mixed[0][0].holo))
            assert np.array_equal(ri, mixed[0][1])
        s = server.stats()
        assert s["device_bytes"] <= budget and s["slots"] == 1


@pytest.mark.parametrize("solver,iterations,holo", [("gp", 5, 128), ("tv", 3, 128), ("gp", 4, 256),
                                                  ("tv", 2, 256)])
def test_batched_slots_give_each_request_its_alone_result(device, qh, solver, iterations, holo):
    """A geometry with batched slots: queued requests of it run together (one front end for
    all their stacks, one graph replay per batch, every solve launch over the batch's volumes)
    and each gets the bits it gets alone on a Pipeline, whatever batch it lands in; distinct
    inputs show the results are routed to their own requests. 64^3 and 128^3 volumes (z round
    trips of 64 and 128 points, FGP rows of 1 and 2 floats per lane)."""
    d = PipelineData(qh, holo=holo, angles=9, iterations=iterations, solver=solver)
    inputs = [(d.holo * np.float32(1.0 + 0.02 * k)).astype(np.float32) for k in range(3)]
    p = qh.Pipeline(d.count, d.size, illuminations=d.illum, geometry=d.geom, **d.kw)
    p.set_background(d.bg)
    alone = [p.run(x) for x in inputs]
    filled = p.filled
    del p
    assert not np.array_equal(alone[0], alone[1])
    server = qh.Server(max_slots=1)
    gid = server.add_geometry(d.count, d.size, illuminations=d.illum, geometry=d.geom,
                              backgrounds=d.bg, max_batch=4, **d.kw)
    order = [0, 1, 2, 2, 0, 1, 1, 2, 0, 0]
    print(f"batched slots {solver}: {len(order)} requests", flush=True)
    ids = [(k, server.submit(gid, inputs[k])) for k in order]
    for k, rid in ids:
        ri, got_filled, latency = server.wait(rid)
        assert np.array_equal(ri, alone[k]), k
        assert got_filled == filled
        assert latency > 0
    s = server.stats()
    assert s["completed"] == len(order) and s["failed"] == 0
    assert 2 <= s["peak_batch"] <= 4
    assert s["batched_requests"] >= 2
    assert s["graph_slots"] == s["slots"]


def test_batched_and_single_geometries_mix(device, qh, mixed):
    """A batched geometry beside unbatched ones, interleaved submissions: every result is the
    request's alone result."""
    server = qh.Server(max_slots=2)
    d0, alone0 = mixed[0]
    g0 = server.add_geometry(d0.count, d0.size, illuminations=d0.illum, geometry=d0.geom,
                             backgrounds=d0.bg, max_batch=3, **d0.kw)
    rest = add_all(server, mixed[1:])
    gids = [g0] + rest
    order = [0, 0, 1, 0, 2, 0, 0, 1, 0]
    ids = [(k, server.submit(gids[k], mixed[k][0].holo)) for k in order]
    for k, rid in ids:
        ri, _, _ = server.wait(rid)
        assert np.array_equal(ri, mixed[k][1]), k
    assert server.stats()["failed"] == 0
