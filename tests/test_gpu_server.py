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
    assert s["slots"] <= max_slots


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
