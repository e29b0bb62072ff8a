"""Benchmark: 256^3 FISTA-TV reconstructions per second on one MI355X.

Workload (the reference README's cell at full scale): 49 off-axis holograms of 512^2 fp32 with
their background holograms -> fields cropped to 256^2 -> ratio, DCT unwrap, Rytov -> Ewald
mapping into 256^3 -> 20 FISTA-TV iterations with 10 FGP dual steps each (TV weight 1e-2) ->
index volume. Synthetic inputs from the package's synthesiser; backgrounds are set once per
geometry (Pipeline.set_background), and each step reconstructs from holograms already in HBM
into a volume in HBM.

Prints one JSON line: throughput (recon/s), per-step latency percentiles, the roofline of the
dominant kernel (the FGP dual step, timed with HIP events on the stream it runs on), a CPU
baseline (the oracle, a C++/OpenMP restatement of the reference's CPU path, timed for one
reconstruction of the same inputs) and the relative error of the HIP volume against it.
"""

from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "256³ FISTA-TV recon/s + p50/p99 latency (served vs plain), 1 MI355X"
HBM_PEAK_GBS = 8000.0
N = 256
ANGLES = 49
ITERATIONS = 20
INNER = 10
TV_WEIGHT = 1e-2


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def fgp_traffic(kernel: str):
    """HBM bytes of the average FGP launch of a prox call measured by rocprofv3 (FETCH_SIZE +
    WRITE_SIZE of each pass, FETCH_SIZE scaled by the copy-kernel calibration, averaged over the
    passes), from the committed summary when it exists."""
    path = ROOT / "profiles" / f"{kernel}_256_hbm_traffic.txt"
    if not path.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", path.read_text())
    return float(m.group(1)) if m else None


def fgp_pass_volumes(inner: int) -> list[int]:
    """Volumes (b, dual components) each FGP launch of one prox call must read and write, in the
    order TvState::prox issues them: two steps per pass, a single last step for odd counts. A
    pass reads b and p_k (3), plus p_{k-1} (3) except on the first step of the call (r = p
    there); it writes p_{k+1} (3, skipped by a last pass of two steps) and p_{k+2} (3)."""
    out = []
    k = 0
    while k < inner:
        first = k == 0
        reads = 1 + 3 + (0 if first else 3)
        if k + 1 < inner:
            last = k + 2 == inner
            out.append(reads + (3 if last else 6))
            k += 2
        else:
            out.append(reads + 3)
            k += 1
    return out


def reconstruction_bytes(count: int, holo: int, n: int, nz: int, iterations: int,
                         inner: int) -> int:
    """Minimum HBM bytes of one FISTA-TV reconstruction on this implementation's stages: every
    stage reads its inputs once and writes its outputs once; a transform counts one read and
    one write of its array however many passes it takes. Front end as the algorithmic model of
    the reference's stages (hologram R2C, crop, inverse, ratio, unwrap, Rytov, 2D FFT); mapping
    with uint16 hit counts; per outer iteration the five transform kernels (momentum fused into
    the first, the data restore into the z pass), the FGP passes and the primal."""
    a, nh, m = count, holo, n
    v = nz * n * n
    h = nz * n * (n // 2 + 1)
    half_holo = a * nh * (nh // 2 + 1)
    radius = 65.0 * n / 256.0
    front = (4 * a * nh * nh + 8 * half_holo            # hologram R2C
             + 8 * a * m * m + 8 * a * np.pi * radius ** 2  # crop
             + 16 * a * m * m                            # field inverse transform
             + 24 * a * m * m                            # ratio, phase, log amplitude
             + 32 * a * m * m                            # unwrap
             + 16 * a * m * m                            # Rytov
             + 16 * a * m * m                            # 2D transform
             + 8 * a * m * m + 10 * v                    # gather and normalise
             + 10 * v + 9 * h                            # Hermitian data and hits
             + 8 * h + 4 * v)                            # initial C2R
    transforms = (8 * v + 8 * h) + 16 * h + 25 * h + 16 * h + (8 * h + 4 * v)
    prox = (sum(fgp_pass_volumes(inner)) + 5) * 4 * v
    return int(front + iterations * (transforms + prox))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cpus() -> dict:
    """Host CPU of the machine running the bench: lscpu's model, sockets and cores, and the
    CPUs this process may run on."""
    out = {"model": cpu_model(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        import subprocess

        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket":
                "cores_per_socket", "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in text.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                v = v.strip()
                out[keys[k.strip()]] = int(v) if v.isdigit() else v
    except (OSError, ValueError):
        pass
    return out


def cpu_baseline(w, runs=3):
    """Times the oracle (the C++/OpenMP restatement of the reference's CPU path) on the same
    inputs for whole reconstructions (median of `runs`), and returns its volume for the parity
    check."""
    import oracle

    g = w.geometry
    geom = (g.wavelength, g.n_medium, g.na, g.pixel)
    cy, cx = (int(np.sign(c) * np.floor(abs(c) + 0.5)) for c in w.center)  # lround
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        ri, _ = oracle.pipeline(w.holograms, w.backgrounds, cy, cx, w.radius, N, w.illuminations,
                                geom, iterations=ITERATIONS, solver="tv", tv_weight=TV_WEIGHT,
                                tv_inner=INNER)
        times.append(time.perf_counter() - t0)
    secs = float(np.median(times))
    log(f"cpu baseline: reconstructions in {', '.join(f'{t:.2f}' for t in times)} s")
    cores = oracle.max_threads()

    # One thread: the front end and 1 and 3 outer iterations, extrapolated to 20 (the outer
    # iterations all do the same work).
    def one_thread(iterations):
        t0 = time.perf_counter()
        oracle.pipeline(w.holograms, w.backgrounds, cy, cx, w.radius, N, w.illuminations, geom,
                        iterations=iterations, solver="tv", tv_weight=TV_WEIGHT, tv_inner=INNER)
        return time.perf_counter() - t0

    oracle.set_threads(1)
    try:
        t1, t3 = one_thread(1), one_thread(3)
    finally:
        oracle.set_threads(cores)
    per_iter = (t3 - t1) / 2.0
    single = (t1 - per_iter) + ITERATIONS * per_iter
    log(f"cpu baseline, one thread: 1 iteration {t1:.2f} s, 3 iterations {t3:.2f} s, "
        f"{ITERATIONS} extrapolated {single:.1f} s")

    # The oracle's own 3D transform against scipy's pocketfft at 256^3 on the same threads.
    x = np.random.default_rng(0).standard_normal((N, N, N)).astype(np.float32)
    fft_ms = {}
    t0 = time.perf_counter()
    oracle.rfft3(x)
    fft_ms["oracle_rfft3"] = round(1000.0 * (time.perf_counter() - t0), 1)
    try:
        import scipy.fft

        scipy.fft.rfftn(x, workers=cores)
        t0 = time.perf_counter()
        scipy.fft.rfftn(x, workers=cores)
        fft_ms["scipy_rfftn"] = round(1000.0 * (time.perf_counter() - t0), 1)
    except ImportError:
        pass
    return {
        "value": round(1.0 / secs, 4),
        "unit": "recon/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"full reconstructions of the same workload by the oracle (C++/OpenMP "
                   f"restatement of the reference CPU path, own FFT), median of {runs}: "
                   f"{secs:.2f} s on {cores} threads of {cpu_model()}"),
        "one_thread": {"value": round(1.0 / single, 5), "unit": "recon/s", "cores": 1,
                       "sample": f"front end + 1 and 3 outer iterations on one thread "
                                 f"({t1:.2f} s, {t3:.2f} s), extrapolated to {ITERATIONS}"},
        "host": host_cpus(),
        "fft_256_ms": fft_ms,
        "reference_published": "Ryzen 5 7500F, 6 cores: FISTA-TV 256^3 in 10,092 ms "
                               "(README.md:454 of the reference), for scale only",
    }, ri


def streaming_reference(torch, mib=256, reps=20):
    """What plain streaming kernels reach on this device, beside the roofline's 8 TB/s: torch's
    device copy (read 1, write 1) and its vectorised addcmul (read 3 float4 streams, write 1) on
    `mib` MiB float32 tensors, timed with events over `reps` launches each."""
    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.full_like(a, 2.0)
    c = torch.full_like(a, 3.0)
    d = torch.empty_like(a)
    out = {"tensor_bytes": n * 4}
    for name, fn, moved in (("copy", lambda: d.copy_(a), 2),
                            ("addcmul", lambda: torch.addcmul(a, b, c, out=d), 4)):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[f"{name}_gbs"] = round(moved * n * 4 / (ms * 1e-3) / 1e9, 1)
    del a, b, c, d
    torch.cuda.empty_cache()
    return out


def fgp_traffic_run(qh, torch, n: int = N) -> int:
    """One TV prox of 10 FGP steps on an n^3 volume (five two-step passes, plus the prox's fill
    and primal kernels) and one copy of a 256 MiB buffer (reads and writes exactly 256 MiB) as
    the counter calibration."""
    from qoraholo_amd import _native

    rng = np.random.default_rng(0)
    b = (rng.standard_normal((n, n, n)) * 0.01).astype(np.float32)
    x = np.empty_like(b)
    _native.check(_native.lib().qh_hip_tv_prox(_native.ptr(b), n, n, n, 0.5, 10, 1,
                                                _native.ptr(x)))
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    # A second calibration with the FGP pass's access pattern: several float4 streams read
    # together (three 268435456-byte inputs) and one written.
    a = torch.full_like(src, 2.0)
    b = torch.full_like(src, 3.0)
    torch.addcmul(src, a, b, out=dst)
    torch.cuda.synchronize()
    log(f"fgp traffic run: 10 FGP steps in 5 passes at {n}^3, a 268435456-byte copy and an "
        "addcmul of three 268435456-byte inputs")
    return 0


BPM_N = 256
BPM_ANGLES = 49
BPM_GEOM = (0.532, 1.337, 1.2, 0.2)


def bpm_bytes(n: int, nz: int, count: int, checkpoint: int) -> tuple[int, int]:
    """Minimum bytes of the fused BPM kernels (forward model; data-term gradient with the
    adjoint recomputing envelopes from checkpoints every `checkpoint` slices, or with every
    envelope stored by the forward when `checkpoint` is 0), per pixel per
    illumination per slice: forward rows 16 (T in, U out) and columns 44 (U, S, step, q in;
    S, T out) = 60; a recomputed slice also stores its envelope (+8); the adjoint columns 44,
    rows 28 (in, envelope, out, gradient term) and the ordered reduction 4. Per slice and not
    per illumination: the screen exp(i phase x) (8 n^2) and the gradient plane (8 n^2)."""
    px = n * n * count * nz
    fwd = 60 * px + 8 * n * n * nz
    recompute = (checkpoint - 1) / checkpoint if checkpoint else 0.0
    grad = (fwd + 8 * px / max(checkpoint, 1) + recompute * (fwd + 8 * px)
            + (44 + 28 + 4) * px + 16 * n * n * nz)
    return int(fwd), int(grad)


def bpm_leg(qh, torch, cpu=True, repeats=5):
    """Beam propagation at 256^2 x 256 with 49 illuminations (NA 0.9): forward model, data-term
    gradient (forward, residual, checkpointed adjoint) and one inversion iteration, on three
    overlapping beads (Delta n <= 0.05). Inputs and outputs of the timed calls stay in HBM
    except for the inversion, which runs from host arrays like the reference's bpm_reconstruct.
    The CPU baseline is the oracle's forward on as many of the 49 illuminations as there are
    threads (one illumination per thread), scaled to 49."""
    from qoraholo_amd import phantom as ph

    g = qh.Geometry(*BPM_GEOM)
    il = qh.circular_scan(BPM_ANGLES, 0.9, g)
    beads = [ph.Sphere(0.0, 0.0, 0.0, 4.0, 1.385), ph.Sphere(2.5, 1.5, -1.0, 3.0, 1.372),
             ph.Sphere(-2.0, -2.0, 2.0, 2.0, 1.360)]
    ri = ph.render_phantom(beads, g, BPM_N, BPM_N)
    x_true = (ri - np.float32(g.n_medium)).astype(np.float32)
    dev = torch.device("cuda:0")
    plan = qh.BpmPlan(BPM_N, BPM_N, il, g)
    xt = torch.from_numpy(x_true).to(dev)
    data = plan.forward(xt)
    x0 = torch.zeros_like(xt)
    torch.cuda.synchronize()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(repeats):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / repeats

    y = torch.empty_like(data)
    grad = torch.empty_like(xt)
    fwd_ms = timed(lambda: plan.forward(xt, out=y))
    grad_ms = timed(lambda: plan.gradient(x0, data, out=grad))
    log(f"bpm: forward {fwd_ms:.2f} ms, gradient {grad_ms:.2f} ms")

    fields = data.cpu().numpy()
    start = np.full_like(ri, np.float32(g.n_medium))
    t0 = time.perf_counter()
    qh.bpm_reconstruct(fields, il, g, start, iterations=1)
    t1 = time.perf_counter()
    _, loss = qh.bpm_reconstruct(fields, il, g, start, iterations=6, return_loss=True)
    t6 = time.perf_counter()
    iter_ms = 1000.0 * ((t6 - t1) - (t1 - t0)) / 5.0
    log(f"bpm: one inversion iteration {iter_ms:.2f} ms (host arrays, from 1 and 6 iterations)")

    a, plane, nz = BPM_ANGLES, BPM_N * BPM_N, BPM_N
    fwd_alg, grad_alg = bpm_bytes(BPM_N, BPM_N, a, plan.checkpoint)
    # The reference's traffic model (88 bytes per pixel per slice per illumination for the
    # forward, two cuFFT transforms and two elementwise passes per slice; 222 GB for a gradient
    # with recomputation from checkpoints every 16 slices).
    fwd_ref = 88 * plane * a * nz
    grad_ref = int(fwd_ref * (1 + 15 / 16) + 100 * plane * a * nz)

    def roof(alg, ref, ms):
        gbs = alg / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "algorithmic_bytes": alg, "achieved": round(gbs, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                "reference_model_bytes": ref,
                "reference_model_frac": round(ref / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    out = {
        "workload": "3 overlapping beads, 256^2 x 256 at 0.2 um, 49 illuminations NA 0.9, "
                    + (f"checkpoint every {plan.checkpoint} slices" if plan.checkpoint
                       else "every envelope stored"),
        "forward_ms": round(fwd_ms, 3),
        "gradient_ms": round(grad_ms, 3),
        "iteration_ms_host_arrays": round(iter_ms, 3),
        "device_bytes": plan.device_bytes,
        "forward_roofline": roof(fwd_alg, fwd_ref, fwd_ms),
        "gradient_roofline": roof(grad_alg, grad_ref, grad_ms),
        "roofline_note": "at 256^2 x 49 each per-slice buffer (S, T: 25.7 MB) stays in the 256 MB "
                         "MALL between the row and column kernels of a slice, so these rates "
                         "are not HBM-bound; the served 384^2 x 81 requests are",
        "loss_6_iterations": [float(v) for v in loss],
        "reference_rtx5070_ms": {"forward_host_in_out": 130, "iteration": 170},
    }
    if cpu:
        import oracle

        cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        sample = max(1, min(a, cores))
        il7 = il[:sample]
        t0 = time.perf_counter()
        want = oracle.bpm_forward(x_true, il7, BPM_GEOM)
        secs = time.perf_counter() - t0
        got = y[:sample].cpu().numpy()
        err = float(np.sqrt(np.sum(np.abs(got.astype(np.complex128) - want) ** 2) /
                            np.sum(np.abs(want.astype(np.complex128) - 1) ** 2)))
        out["cpu_baseline"] = {"value": round(secs * 1000.0 * a / sample, 1), "unit": "ms/forward",
                               "cores": cores, "kind": "port",
                               "sample": f"oracle forward on {sample} of {a} illuminations, "
                                         f"{secs:.2f} s, scaled to {a}"}
        out["parity"] = {"scattered_rel_l2_vs_oracle": err, "illuminations": sample}
        log(f"bpm: oracle forward on {sample} illuminations {secs:.2f} s; parity {err:.2e}")
    return out


MIXED_BATCH = 4
# Reconstructions in flight in the timed throughput loop (2: 50.0 against 49.1 recon/s for one,
# 3: 49.2, measured).
IN_FLIGHT = 2


def served_mixed(qh, rates=(0.25, 0.5, 1.0, 2.0), per_rate=100, seed=0, max_slots=8):
    """Config 3: a served stream of mixed shapes and solvers at Poisson rates set against the
    plain path's measured capacity. Shapes uniform over 25 x 256^2 -> 128^3, 49 x 512^2 -> 256^3
    and 81 x 768^2 -> 384^3 (random cells of 3-8 nested or disjoint spheres), solvers uniform
    over GP 20 and FISTA-TV 20 x 10. Capacity = 1 / mean service time of the plain per-call
    chain over the six kinds (median of 3 timed requests each); also reported: serial
    Pipeline.run per kind. Up to eight requests in flight (four served as many requests per
    second at twice the plain capacity, with a longer queue); idle slots stay resident."""
    from qoraholo_amd import workloads

    kinds = []
    t0 = time.perf_counter()
    for i, (angles, n) in enumerate(((25, 128), (49, 256), (81, 384))):
        w = workloads.mixed_cell(n, angles, seed=seed + i)
        for name, kw in (("gp", dict(iterations=ITERATIONS, solver="gp")),
                         ("tv", dict(iterations=ITERATIONS, solver="tv", tv_weight=TV_WEIGHT,
                                     tv_inner=INNER))):
            kinds.append((f"{n}^3 {name}", w, kw))
    log(f"mixed stream: workloads in {time.perf_counter() - t0:.1f} s")

    def plain_once(w, kw):
        crop = dict(center=w.center, radius=w.radius, output_size=w.field_size)
        ratio = qh.field_ratio(qh.extract_fields(w.holograms, **crop),
                               qh.extract_fields(w.backgrounds, **crop))
        return qh.reconstruct(qh.rytov(ratio), w.illuminations, w.geometry, **kw)

    def median_ms(fn, runs=3):
        times = []
        for _ in range(runs):
            s0 = time.perf_counter()
            fn()
            times.append(1000.0 * (time.perf_counter() - s0))
        return float(np.median(times))

    plain_ms, pipe_ms = {}, {}
    for name, w, kw in kinds:
        plain_once(w, kw)
        plain_ms[name] = median_ms(lambda: plain_once(w, kw))
        p = qh.Pipeline(w.holograms.shape[0], w.holograms.shape[1], w.center, w.illuminations,
                        w.geometry, radius=w.radius, output_size=w.field_size, **kw)
        p.set_background(w.backgrounds)
        p.run(w.holograms)
        pipe_ms[name] = median_ms(lambda: p.run(w.holograms))
        del p
    capacity = 1000.0 / float(np.mean(list(plain_ms.values())))
    log(f"mixed stream: plain ms {plain_ms}, pipeline ms {pipe_ms}, capacity {capacity:.2f}/s")

    server = qh.Server(max_slots=max_slots)
    # The 128^3 kinds (about 1.6 ms of device time each) get batched slots: a slot takes up to
    # four queued requests of its kind at once; the larger kinds are device-bound and run alone.
    gids = [server.add_geometry(w.holograms.shape[0], w.holograms.shape[1], w.center,
                                w.illuminations, w.geometry, w.backgrounds, radius=w.radius,
                                output_size=w.field_size,
                                max_batch=MIXED_BATCH if w.holograms.shape[1] <= 256 else 1, **kw)
            for _, w, kw in kinds]
    for k, (_, w, _) in enumerate(kinds):
        server.wait(server.submit(gids[k], w.holograms))
    rng = np.random.default_rng(seed)
    out = {"kinds": [k[0] for k in kinds], "max_batch_128": MIXED_BATCH,
           "plain_ms": {k: round(v, 2) for k, v in plain_ms.items()},
           "pipeline_ms": {k: round(v, 2) for k, v in pipe_ms.items()},
           "plain_capacity_per_s": round(capacity, 3), "rates": []}
    for factor in rates:
        rate = factor * capacity
        server.reset_stats()
        order = rng.integers(0, len(kinds), per_rate)
        gaps = rng.exponential(1.0 / rate, per_rate)
        ids, peak_queue = [], 0
        tick = time.perf_counter()
        for k, gap in zip(order, gaps):
            tick += gap
            delay = tick - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(server.submit(gids[int(k)], kinds[int(k)][1].holograms))
            peak_queue = max(peak_queue, server.stats()["queued"])
        for rid in ids:
            server.wait(rid)
        st = server.stats()
        row = {"rate_x_capacity": factor, "rate_per_s": round(rate, 3),
               "recon_s": round(st["recon_per_s"], 3), "p50_ms": round(st["p50_ms"], 2),
               "p99_ms": round(st["p99_ms"], 2), "requests": per_rate, "admitted": st["completed"],
               "peak_queued": peak_queue, "slots": st["slots"],
               "batched_requests": st["batched_requests"], "peak_batch": st["peak_batch"],
               "device_gb": round(st["device_bytes"] / 1e9, 2)}
        out["rates"].append(row)
        log(f"mixed stream at {factor}x: {row}")
    return out


def served_bpm(qh, sizes=((192, 49), (384, 81)), iterations=40, seed=0, repeats=None,
               max_slots=4, stream=True, poisson_load=None):
    """Config 5: BPM inversion requests from host arrays. Each request is the reference's recipe
    (examples/bpm_figure.py:74-88): Rytov of the field ratios, 20 GP iterations for the start,
    then 40 full-batch FISTA iterations of the BPM with the TV prox (weight 1e-2, 10 inner
    steps). Sizes 192^2 x 192 with 49 illuminations and 384^2 x 384 with 81; random cells at
    0.2 um. First each size through the per-call functions, one at a time (the stage split and
    the per-iteration time; per iteration, the minimum bytes of a gradient at the checkpoint
    interval the model resolves to are the algorithmic figure, beside the reference's traffic
    model -- forward, 15/16 recomputation from checkpoints, adjoint); then
    the same requests through the Server, its BPM geometries resident and admitted by their
    measured HBM footprint: a burst of `repeats[n]` requests of side n (4 of 192, 2 of 384),
    interleaved, with inversions/s, p50/p99 latency and each result compared with the per-call
    one. With `poisson_load`, the requests arrive instead by a Poisson process at that fraction
    of the per-call capacity (1 / mean per-call request time), so that the latency percentiles
    are those of a loaded server rather than of a queue drained from a burst."""
    from qoraholo_amd import workloads

    out = []
    reqs = []
    for n, a in sizes:
        t0 = time.perf_counter()
        req = workloads.bpm_request(n, a, seed=seed)
        log(f"served bpm {n}^3 x {a}: request built in {time.perf_counter() - t0:.1f} s")
        g, il, f = req.geometry, req.illuminations, req.fields
        flat = np.full((n, n, n), np.float32(g.n_medium))
        qh.bpm_reconstruct(f, il, g, flat, iterations=2, tv_weight=TV_WEIGHT)  # warm-up
        s0 = time.perf_counter()
        ry = qh.rytov(f)
        s1 = time.perf_counter()
        start = qh.reconstruct(ry, il, g, iterations=ITERATIONS)
        s2 = time.perf_counter()
        ri, loss = qh.bpm_reconstruct(f, il, g, start, iterations=iterations,
                                      tv_weight=TV_WEIGHT, tv_inner=INNER, return_loss=True)
        s3 = time.perf_counter()
        # One more solve of 2 iterations separates the per-iteration cost from the fixed one.
        qh.bpm_reconstruct(f, il, g, start, iterations=2, tv_weight=TV_WEIGHT, tv_inner=INNER)
        s4 = time.perf_counter()
        per_iter = 1000.0 * ((s3 - s2) - (s4 - s3)) / (iterations - 2)
        # The per-call functions resolve the checkpoint interval from the free memory as a
        # plan built now does (0 when every envelope of the illuminations fits).
        checkpoint = qh.BpmPlan(n, n, il, g).checkpoint
        grad_alg = bpm_bytes(n, n, a, checkpoint)[1]
        grad_ref = int((88 * (1 + 15 / 16) + 100) * n * n * a * n)
        truth = req.truth.astype(np.float64)
        row = {
            "size": n,
            "workload": f"random cell {n}^2 x {n} at 0.2 um, {a} illuminations NA 0.9, "
                        f"Rytov + GP {ITERATIONS} start, {iterations} FISTA iterations, "
                        f"TV {TV_WEIGHT} x {INNER}",
            "request_ms": round(1000.0 * (s3 - s0), 1),
            "rytov_ms": round(1000.0 * (s1 - s0), 1),
            "gp_start_ms": round(1000.0 * (s2 - s1), 1),
            "bpm_ms": round(1000.0 * (s3 - s2), 1),
            "iteration_ms": round(per_iter, 2),
            "checkpoint": checkpoint,
            "iteration_roofline": {"bound": "hbm", "algorithmic_bytes": grad_alg,
                                   "achieved": round(grad_alg / (per_iter * 1e-3) / 1e9, 1),
                                   "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": round(grad_alg / (per_iter * 1e-3) / 1e9
                                                 / HBM_PEAK_GBS, 4),
                                   "reference_model_bytes": grad_ref,
                                   "reference_model_frac": round(
                                       grad_ref / (per_iter * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                   "note": "gradient bytes only; the iteration also runs the "
                                           "TV prox and the FISTA steps"},
            "loss_first_last": [float(loss[0]), float(loss[-1])],
            "rmse_start_end": [float(np.sqrt(np.mean((start - truth) ** 2))),
                               float(np.sqrt(np.mean((ri - truth) ** 2)))],
        }
        log(f"served bpm {n}^3 x {a}: {row}")
        out.append(row)
        reqs.append((req, ri, loss))
    if not stream:
        return {"requests": out}

    server = qh.Server(max_slots=max_slots)
    gids = [server.add_bpm(n, req.illuminations, req.geometry, start_iterations=ITERATIONS,
                           iterations=iterations, tv_weight=TV_WEIGHT, tv_inner=INNER)
            for (n, _), (req, _, _) in zip(sizes, reqs)]
    for k, gid in enumerate(gids):  # one untimed request per geometry
        server.wait(server.submit_bpm(gid, reqs[k][0].fields))
    server.reset_stats()
    order = []
    repeats = repeats or {192: 4, 384: 2}
    counts = [repeats.get(n, 2) for n, _ in sizes]
    for i in range(max(counts)):
        order += [k for k in range(len(sizes)) if i < counts[k]]
    log(f"served bpm stream: {len(order)} requests, sizes {[sizes[k][0] for k in order]}")
    plain_s = sum(out[k]["request_ms"] for k in order) / 1000.0

    def run(sequence, rate=None):
        """Submits the requests of `sequence` (a burst, or Poisson arrivals at `rate` per second)
        and waits for all; returns (stats, wall seconds, every result equal to its per-call one)."""
        server.reset_stats()
        t0 = time.perf_counter()
        if rate:
            gaps = np.random.default_rng(seed).exponential(1.0 / rate, len(sequence))
            ids = []
            tick = time.perf_counter()
            for k, gap in zip(sequence, gaps):
                tick += gap
                delay = tick - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids.append((k, server.submit_bpm(gids[k], reqs[k][0].fields)))
        else:
            ids = [(k, server.submit_bpm(gids[k], reqs[k][0].fields)) for k in sequence]
        same = True
        for k, rid in ids:
            ri, loss, _ = server.wait(rid)
            same = (same and bool(np.array_equal(ri, reqs[k][1]))
                    and bool(np.array_equal(loss, reqs[k][2])))
        return server.stats(), time.perf_counter() - t0, same

    per_call = len(order) / plain_s
    if poisson_load:
        # Capacity from a burst of 16, then the latency of a loaded server: Poisson arrivals at
        # `poisson_load` times that capacity.
        burst = order[:16]
        st, wall, same_b = run(burst)
        capacity = st["recon_per_s"]
        log(f"served bpm burst: {len(burst)} requests, {capacity:.4f}/s")
        st, wall, same = run(order, poisson_load * capacity)
        same = same and same_b
        stream = {"requests": len(order), "sizes": sorted(set(sizes[k][0] for k in order)),
                  "burst_requests": len(burst),
                  "burst_inversions_per_s": round(capacity, 4),
                  "arrivals": f"Poisson at {poisson_load} x the burst capacity "
                              f"({poisson_load * capacity:.3f}/s)",
                  "completed_per_s": round(st["recon_per_s"], 4)}
    else:
        st, wall, same = run(order)
        stream = {"requests": len(order), "sizes": [sizes[k][0] for k in order],
                  "arrivals": "burst", "inversions_per_s": round(st["recon_per_s"], 4)}
    stream.update({"wall_s": round(wall, 2),
                   "p50_ms": round(st["p50_ms"], 1), "p99_ms": round(st["p99_ms"], 1),
                   "per_call_inversions_per_s": round(per_call, 4),
                   "slots": st["slots"], "device_bytes": st["device_bytes"],
                   "budget_bytes": st["budget_bytes"],
                   "results_equal_per_call_bitwise": same})
    log(f"served bpm stream: {stream}")
    return {"requests": out, "served_stream": stream}


def pct(v, q):
    return round(float(np.percentile(v, q)), 3)


def serving_comparison(qh, w, plain_requests=100, serial_requests=100, served_requests=100,
                       seed=0):
    """Requests from host memory to host memory on the same workload:
    - plain: the per-call chain a user of the reference's API runs (extract_fields for sample
      and background, field_ratio on the host, rytov, reconstruct), serially;
    - pipeline: Pipeline.run, serially (background fields resident);
    - served: the Server with two slots, a burst of requests submitted at once (throughput),
      then a Poisson stream at 80% of that throughput (latency under load)."""
    g = w.geometry
    kw = dict(iterations=ITERATIONS, solver="tv", tv_weight=TV_WEIGHT, tv_inner=INNER)
    crop = dict(center=w.center, radius=w.radius, output_size=N)
    out = {}

    def plain_once():
        ratio = qh.field_ratio(qh.extract_fields(w.holograms, **crop),
                               qh.extract_fields(w.backgrounds, **crop))
        return qh.reconstruct(qh.rytov(ratio), w.illuminations, g, **kw)

    plain_once()
    lat = []
    t0 = time.perf_counter()
    for _ in range(plain_requests):
        s0 = time.perf_counter()
        ri_plain = plain_once()
        lat.append(1000.0 * (time.perf_counter() - s0))
    out["plain"] = {"recon_s": round(plain_requests / (time.perf_counter() - t0), 3),
                    "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "requests": plain_requests}
    log(f"plain per-call: {out['plain']}")

    p = qh.Pipeline(ANGLES, 2 * N, w.center, w.illuminations, g, radius=w.radius, output_size=N,
                    **kw)
    p.set_background(w.backgrounds)
    p.run(w.holograms)
    lat = []
    t0 = time.perf_counter()
    for _ in range(serial_requests):
        s0 = time.perf_counter()
        ri_pipe = p.run(w.holograms)
        lat.append(1000.0 * (time.perf_counter() - s0))
    out["pipeline"] = {"recon_s": round(serial_requests / (time.perf_counter() - t0), 3),
                       "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                       "requests": serial_requests}
    log(f"serial pipeline: {out['pipeline']}")
    del p

    # Two slots: one computes while the other moves its holograms in and its volume out; more
    # slots of the same geometry only share the device between more FGP passes.
    server = qh.Server(max_slots=2)
    gid = server.add_geometry(ANGLES, 2 * N, w.center, w.illuminations, g, w.backgrounds,
                              radius=w.radius, output_size=N, **kw)
    for rid in [server.submit(gid, w.holograms) for _ in range(2)]:
        server.wait(rid)
    server.reset_stats()
    t0 = time.perf_counter()
    ids = [server.submit(gid, w.holograms) for _ in range(served_requests)]
    results = [server.wait(rid) for rid in ids]
    burst_s = served_requests / (time.perf_counter() - t0)
    st = server.stats()
    ri_served = results[-1][0]
    out["served_burst"] = {"recon_s": round(burst_s, 3), "p50_ms": round(st["p50_ms"], 3),
                           "p99_ms": round(st["p99_ms"], 3), "requests": served_requests,
                           "slots": st["slots"], "device_bytes": st["device_bytes"]}
    log(f"served burst: {out['served_burst']}")

    rate = 0.8 * burst_s
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, served_requests)
    server.reset_stats()
    ids = []
    t0 = time.perf_counter()
    for gap in gaps:
        t0 += gap
        delay = t0 - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(server.submit(gid, w.holograms))
    for rid in ids:
        server.wait(rid)
    st = server.stats()
    out["served_poisson"] = {"rate_per_s": round(rate, 3), "recon_s": round(st["recon_per_s"], 3),
                             "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                             "requests": served_requests}
    log(f"served poisson: {out['served_poisson']}")
    out["served_vs_plain"] = round(out["served_burst"]["recon_s"] / out["plain"]["recon_s"], 3)
    out["served_equals_pipeline_bitwise"] = bool(np.array_equal(ri_served, ri_pipe))
    nm = g.n_medium
    out["plain_vs_served_contrast_rel_l2"] = float(
        np.sqrt(np.sum((ri_plain.astype(np.float64) - ri_served) ** 2) /
                np.sum((ri_served.astype(np.float64) - nm) ** 2)))
    return out, ri_served


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the oracle run (CPU baseline and full-size parity)")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the served-versus-plain comparison (host-memory requests)")
    ap.add_argument("--no-bpm", action="store_true", help="skip the beam propagation leg")
    ap.add_argument("--bpm-only", action="store_true",
                    help="run only the beam propagation leg and print its JSON object")
    ap.add_argument("--served-mixed", action="store_true",
                    help="run only the mixed served stream (128^3/256^3/384^3, GP and TV, "
                         "Poisson arrivals at 0.25x-2x the plain capacity) and print its JSON")
    ap.add_argument("--served-bpm", action="store_true",
                    help="only the served BPM inversion requests (192^3 x 49, 384^3 x 81)")
    ap.add_argument("--served-bpm-size", type=int, choices=(192, 384), default=0,
                    help="with --served-bpm: only the request of this size")
    ap.add_argument("--no-served-stream", action="store_true",
                    help="with --served-bpm: only the per-call requests, not the served stream")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed served stream (configuration 3) in the default run")
    ap.add_argument("--no-served-bpm", action="store_true",
                    help="skip the served 192^3 x 49 BPM inversions (configuration 5) in the "
                         "default run")
    ap.add_argument("--in-flight", type=int, default=IN_FLIGHT, choices=(1, 2),
                    help="reconstructions in flight in the timed throughput loop (1 for kernel "
                         "traces whose durations must not overlap)")
    ap.add_argument("--fgp-traffic", action="store_true",
                    help="run only one TV prox (10 FGP steps in 5 passes) and two calibration "
                         "kernels, for the rocprofv3 FETCH_SIZE / WRITE_SIZE runs that measure "
                         "HBM traffic")
    ap.add_argument("--fgp-traffic-size", type=int, default=N,
                    help="volume side of the --fgp-traffic run (default 256)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the reconstruction path runs on one device and the "
              f"multi-GPU design (requests sharded across devices) is not built; got --gpus "
              f"{args.gpus}", file=sys.stderr)
        return 2

    import torch

    import qoraholo_amd as qh
    from qoraholo_amd import workloads

    if args.fgp_traffic:
        return fgp_traffic_run(qh, torch, args.fgp_traffic_size)

    if not qh.hip_available():
        raise SystemExit("no HIP device or libqoraholo_hip.so missing")
    if args.served_mixed:
        print(json.dumps(served_mixed(qh)), flush=True)
        return 0
    if args.served_bpm:
        sizes = ((192, 49), (384, 81))
        if args.served_bpm_size:
            sizes = tuple(v for v in sizes if v[0] == args.served_bpm_size)
        print(json.dumps(served_bpm(qh, sizes=sizes, stream=not args.no_served_stream)), flush=True)
        return 0
    if args.bpm_only:
        print(json.dumps(bpm_leg(qh, torch, cpu=not args.no_cpu_baseline, repeats=args.steps)),
              flush=True)
        return 0
    t0 = time.perf_counter()
    w = workloads.readme_cell(n=N, angles=ANGLES, seed=0)
    log(f"workload generated in {time.perf_counter() - t0:.1f} s")

    p = qh.Pipeline(ANGLES, 2 * N, w.center, w.illuminations, w.geometry, radius=w.radius,
                    output_size=N, iterations=ITERATIONS, solver="tv", tv_weight=TV_WEIGHT,
                    tv_inner=INNER)
    p.set_background(w.backgrounds)
    dev = torch.device("cuda:0")
    holo = torch.from_numpy(w.holograms).to(dev)
    out = torch.empty((N, N, N), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    # A second pipeline of the same geometry on a stream of its own, so that two reconstructions
    # are in flight: each fills the other's kernel tails (the served loop does the same with two
    # slots). Both streams at the default priority.
    in_flight = max(1, min(2, args.in_flight))
    pipes, outs = [p], [out]
    if in_flight == 2:
        p2 = qh.Pipeline(ANGLES, 2 * N, w.center, w.illuminations, w.geometry, radius=w.radius,
                         output_size=N, iterations=ITERATIONS, solver="tv", tv_weight=TV_WEIGHT,
                         tv_inner=INNER)
        p2.set_background(w.backgrounds)
        pipes.append(p2)
        outs.append(torch.empty_like(out))
    streams = [torch.cuda.Stream() for _ in pipes]
    for _ in range(args.warmup):
        for k in range(in_flight):
            pipes[k].run_device(holo, outs[k], stream=streams[k])
        p.run_device(holo, out)
    torch.cuda.synchronize()
    log(f"warmup done ({args.warmup} steps)")

    # Throughput: K reconstructions, each replayed as one captured hipGraph, in_flight at a
    # time (step i on pipeline and stream i mod in_flight).
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        k = i % in_flight
        pipes[k].run_device(holo, outs[k], stream=streams[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    log(f"timed {args.steps} steps, {in_flight} in flight: {ms_per_step:.2f} ms per "
        "reconstruction (graph replays)")
    in_flight_equal = all(bool(torch.equal(out, o)) for o in outs[1:])
    # The same K back to back on one stream.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        p.run_device(holo, out)
    torch.cuda.synchronize()
    ms_one = 1000.0 * (time.perf_counter() - t0) / args.steps
    log(f"timed {args.steps} steps, one in flight: {ms_one:.2f} ms per reconstruction")
    del pipes, outs

    # FGP kernel durations: K more reconstructions launched kernel by kernel with HIP events
    # around every prox on the pipeline's stream (events inside a graph are not time-stamped).
    p.time_fgp(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        p.run_device(holo, out)
    torch.cuda.synchronize()
    ms_kernels = 1000.0 * (time.perf_counter() - t0) / args.steps
    t = p.timings
    p.time_fgp(False)
    log(f"timed {args.steps} steps kernel by kernel: {ms_kernels:.2f} ms per reconstruction")

    # Latency of the plain path: each request submitted and waited for on its own.
    lat = []
    for _ in range(args.steps):
        s0 = time.perf_counter()
        p.run_device(holo, out)
        torch.cuda.synchronize()
        lat.append(1000.0 * (time.perf_counter() - s0))
    p50, p99 = np.percentile(lat, [50, 99])

    vol_bytes = N * N * N * 4
    launches = max(1, t["fgp_launches"])
    steps_per_launch = t["fgp_steps"] / launches
    kernel = "fgp_pair" if steps_per_launch == 2 else "fgp_step"
    passes = fgp_pass_volumes(INNER)
    assert len(passes) * ITERATIONS == launches, (passes, launches)
    # What the launches must move: per pass b and the dual iterates it reads, the iterates it
    # writes (13 volumes for a two-step pass inside a prox call, 10 for its first and last).
    alg_bytes = ITERATIONS * sum(passes) * vol_bytes
    fgp_avg_ms = t["fgp_ms"] / launches
    per_launch = alg_bytes / launches
    achieved = per_launch / (fgp_avg_ms * 1e-3) / 1e9
    # The reference's model: 13 volumes per step (read b, r(3), p(3); write p(3), r(3)). A pass
    # of two steps moves half of that, so its rate on this model can exceed the peak.
    ref_bytes = int(13 * vol_bytes * steps_per_launch)
    roofline = {
        "bound": "hbm",
        "kernel": f"k_{kernel}",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": fgp_traffic(kernel),
        "launch_ms": round(fgp_avg_ms, 5),
        "steps_per_launch": steps_per_launch,
        "algorithmic_bytes": int(per_launch),
        "volumes_per_pass": passes,
        "reference_model_bytes": ref_bytes,
        "reference_model_gbs": round(ref_bytes / (fgp_avg_ms * 1e-3) / 1e9, 1),
        "reference_model_frac": round(ref_bytes / (fgp_avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }
    if roofline["traffic"]:
        # The physical rate: measured HBM bytes of the average launch of a prox call (the mean
        # over its passes, which move 10, 13, 13, 13 and 10 volumes) over the mean launch
        # duration of the same passes.
        roofline["traffic_gbs"] = round(roofline["traffic"] / (fgp_avg_ms * 1e-3) / 1e9, 1)
        roofline["traffic_frac"] = round(roofline["traffic_gbs"] / HBM_PEAK_GBS, 4)
    roofline["streaming_reference"] = streaming_reference(torch)
    if roofline.get("traffic_gbs"):
        # The measured traffic rate against what plain streaming kernels reach on this device in
        # this run (the peak in `frac` is the datasheet's).
        ref = roofline["streaming_reference"]
        roofline["traffic_vs_copy"] = round(roofline["traffic_gbs"] / ref["copy_gbs"], 4)
        roofline["traffic_vs_addcmul"] = round(roofline["traffic_gbs"] / ref["addcmul_gbs"], 4)
    recon_bytes = reconstruction_bytes(ANGLES, 2 * N, N, N, ITERATIONS, INNER)
    per_recon = {
        "bytes": recon_bytes,
        "ms": round(ms_per_step, 3),
        "achieved": round(recon_bytes / (ms_per_step * 1e-3) / 1e9, 1),
        "unit": "GB/s",
        "frac": round(recon_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "model": "each kernel's own minimum bytes (inputs read once, outputs written once), "
                 "summed over one reconstruction",
    }
    stages = {k: round(v, 3) for k, v in t.items() if k not in ("fgp_steps",)}

    result = {
        "metric": METRIC,
        "value": round(1000.0 / ms_per_step, 3),
        "unit": "recon/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("readme_cell: 49 x 512^2 fp32 holograms + backgrounds -> 256^2 fields -> "
                         "256^3 FISTA-TV 20x10, tv_weight 1e-2, inputs resident in HBM"),
            "holograms": [ANGLES, 2 * N, 2 * N],
            "volume": [N, N, N],
            "solver": f"fista-tv {ITERATIONS}x{INNER}",
            "launch": "one captured hipGraph per reconstruction",
            "in_flight": in_flight,
            "value_is": f"device-resident throughput (holograms already in HBM, volume left in "
                        f"HBM), {in_flight} reconstruction(s) in flight on their own streams; "
                        f"one in flight: ms_per_step_one_in_flight; host-array rates: "
                        f"serving.pipeline (serial Pipeline.run) and serving.served_burst",
        },
        "ms_per_step_one_in_flight": round(ms_one, 3),
        "in_flight_results_equal_bitwise": in_flight_equal,
        "ms_per_step_kernel_by_kernel": round(ms_kernels, 3),
        "latency_ms": {"device_resident": {"p50": round(float(p50), 3),
                                           "p99": round(float(p99), 3)}},
        "stage_ms": stages,
        "roofline": roofline,
        "per_reconstruction": per_recon,
    }

    if not args.no_serving:
        serving, _ = serving_comparison(qh, w)
        result["serving"] = serving
        result["latency_ms"]["plain"] = {"p50": serving["plain"]["p50_ms"],
                                         "p99": serving["plain"]["p99_ms"]}
        result["latency_ms"]["served"] = {"p50": serving["served_poisson"]["p50_ms"],
                                          "p99": serving["served_poisson"]["p99_ms"]}

    if not args.no_mixed:
        # Configuration 3 at 1x and 2x the plain path's capacity, 100 Poisson arrivals each.
        result["served_mixed"] = served_mixed(qh, rates=(1.0, 2.0), per_rate=100)

    if not args.no_bpm:
        result["bpm"] = bpm_leg(qh, torch, cpu=not args.no_cpu_baseline)

    if not args.no_served_bpm:
        # Configuration 5's 192^2 x 192 / 49-illumination requests: 100 Poisson arrivals at 0.8x
        # the per-call capacity (the 384^3 / 81 requests take 9 s each: --served-bpm).
        result["served_bpm"] = served_bpm(qh, sizes=((192, 49),), repeats={192: 100},
                                          poisson_load=0.8)

    if not args.no_cpu_baseline:
        result["cpu_baseline"], want = cpu_baseline(w)
        got = out.cpu().numpy().astype(np.float64)
        nm = w.geometry.n_medium
        err = float(np.sqrt(np.sum((got - want) ** 2) / np.sum((want.astype(np.float64) - nm) ** 2)))
        raw = float(np.sqrt(np.sum((got - want) ** 2) / np.sum(want.astype(np.float64) ** 2)))
        result["parity"] = {"contrast_rel_l2_vs_oracle": err, "ri_rel_l2_vs_oracle": raw,
                            "tolerance": 5e-5}
        # The same reconstruction stated against the oracle in double precision: the device and
        # the float oracle (the reference's arithmetic) are about equally far from it.
        import oracle

        g = w.geometry
        cy, cx = (int(np.sign(c) * np.floor(abs(c) + 0.5)) for c in w.center)
        t0 = time.perf_counter()
        want64, _ = oracle.pipeline(w.holograms, w.backgrounds, cy, cx, w.radius, N,
                                    w.illuminations, (g.wavelength, g.n_medium, g.na, g.pixel),
                                    dtype=np.float64, iterations=ITERATIONS, solver="tv",
                                    tv_weight=TV_WEIGHT, tv_inner=INNER)
        log(f"fp64 oracle: one reconstruction in {time.perf_counter() - t0:.2f} s")
        c64 = np.sum((want64 - nm) ** 2)
        result["parity"]["contrast_rel_l2_vs_fp64"] = float(np.sqrt(np.sum((got - want64) ** 2) / c64))
        result["parity"]["oracle_float_contrast_rel_l2_vs_fp64"] = float(
            np.sqrt(np.sum((want.astype(np.float64) - want64) ** 2) / c64))
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
