#!/usr/bin/env python3
"""Headline benchmark: 2D Delaunay triangulation + full alpha-shape spectrum on one MI355X.

Workload (the plain method at scale): 1,048,576 i.i.d. U[0,1)^2 FP64 points (seed 2), PBA
texture 4096 x 4096.  A step is one complete pass of the hot path over that set: discrete
Voronoi diagram, reconstruction, shifting, missing-site insertion, boundary removal, flipping,
canonical output, then rho, attachment, exact spectrum ranks, intervals and the master list.
Inputs are resident in HBM when the timed region starts; outputs stay on the device.

Prints one JSON line.  `roofline` covers the dominant hand-written kernel (the largest share of
GPU time after rocPRIM's sort passes, timed with HIP events on the stream it runs on, through the
library's kernel probe); `cpu_baseline` is the sequential CPU oracle on the same set, one core.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_POINTS = 1048576
TEXTURE = 4096
SEED = 2
HBM_PEAK_GBS = 8000.0


# Algorithmic (compulsory) bytes of one run of a kernel, from the work counters of that run.
ALGO_BYTES = {
    "k_edge_intervals": {
        "model": "8N + 24E + 4F: the interval half of the intervals + master-list stage (vertex "
                 "interval pair 8 B per point; rank, lower and upper interval ends 24 B per edge read "
                 "or written once; triangle rank 4 B)",
        "bytes": lambda st, T, n: 8.0 * n + 24.0 * st["nEdges"] + 4.0 * st["nTris"]},
    "k_frame_remove": {
        "model": "12 B per frame vertex examined (state 4 + priority 8) + 88 B per fan triangle of a "
                 "removed frame vertex (triangle record 32 read, chain position 16 read, fill or dead "
                 "slot 32 written, anchor and back-link 8 written)",
        "bytes": lambda st, T, n: 12.0 * st["frameSeen"] + 88.0 * st["frameChain"]},
    "k_flip_test": {
        "model": "32 B per listed triangle: the flipping stage's 24F + 16V per pass over F triangles "
                 "and V = F / 2 vertices (24 B vertex + neighbour record, 16 B position); the "
                 "triangles listed by the grid passes (the single-workgroup tail's are not k_flip_test's)",
        "bytes": lambda st, T, n: 32.0 * (st["flipListed"] - st["flipTailListed"]),
        "gather_model": "156 B per listed triangle (list entry 4, vertices + links 32, three neighbour "
                        "stamps and far vertices 24, six positions 96)",
        "gather_bytes": lambda st, T, n: 156.0 * (st["flipListed"] - st["flipTailListed"])},
    "k_pba_color": {
        "model": "8 B per pixel (envelope read 4 + label written 4)",
        "bytes": lambda st, T, n: 8.0 * T * T * st["pbaIterations"]},
}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary (FETCH_SIZE / WRITE_SIZE
    passes scaled by a 256 MiB copy), or None when it has no row for the kernel."""
    path = os.path.join(ROOT, "profiles", "c2_hbm_traffic.txt")
    try:
        for line in open(path):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 5 and cells[0] == f"`{kernel}`":
                return round(float(cells[4]) * 1e6)
    except OSError:
        pass
    return None


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        return next((l.split(":", 1)[1].strip() for l in out.splitlines() if l.startswith("Model name")), "")
    except Exception:
        return ""


def cpu_baseline(n, seed, runs=5):
    """The oracle (sequential C: Shewchuk predicates, Bowyer-Watson DT, exact alpha ranks) on the
    full workload, one core: one warm-up run, then the median of `runs` timed runs (about 3 s
    each).  Also returns the last run's triangles and alpha arrays for the parity check."""
    import numpy as np
    from oracle import pyoracle as O
    xy = O.gen_uniform(seed, n)
    times = []
    for r in range(runs + 1):
        t0 = time.perf_counter()
        tris, _ = O.delaunay(xy)
        alpha = O.alpha(xy, tris)
        if r:
            times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    base = {"value": round(n / med / 1e6, 4), "unit": "Mpoints/s", "cores": 1, "kind": "port",
            "sample": f"full workload: {n} uniform points seed {seed}, DT + alpha spectrum, median of "
                      f"{runs} runs after one warm-up ({min(times):.2f}-{max(times):.2f} s) on one core of "
                      f"{cpu_model() or 'the host'} (nproc={os.cpu_count()})"}
    return base, tris, alpha


def oracle_check(gpu_tris, gpu_alpha, ref_tris, ref_alpha):
    """The device result of the timed workload against the oracle run of the CPU baseline:
    canonical triangles, spectrum values in ulps, and every rank array."""
    import numpy as np
    out = {"triangles_equal": bool(np.array_equal(gpu_tris, ref_tris))}
    g, r = np.asarray(gpu_alpha["spectrum"]), np.asarray(ref_alpha["spectrum"])
    if g.shape == r.shape:
        ulp = np.abs(g.view(np.int64) - r.view(np.int64))
        out["rho_max_ulp"] = int(ulp.max()) if ulp.size else 0
    else:
        out["rho_max_ulp"] = None
        out["spectrum_sizes"] = [int(g.size), int(r.size)]
    ranks = ("tri_ro_rank", "edge_t", "edge_vopp", "edge_ro_rank", "edge_nu_min", "edge_nu_max", "vert_nu_min",
             "vert_nu_max", "master_simplex", "master_st", "master_at", "spec_pointers")
    out["ranks_equal"] = all(np.array_equal(np.asarray(gpu_alpha[k]), np.asarray(ref_alpha[k])) for k in ranks)
    return out


def cpu_batched(nsets, sample, seed=4):
    """The oracle on the first `sample` sets of the batched workload, one set per task on up to 16
    host threads (the oracle is C called through ctypes, which releases the GIL, so the threads
    run in parallel; no child process is started)."""
    from concurrent.futures import ThreadPoolExecutor
    from galpha2d_amd import datagen
    from oracle import pyoracle as O
    sets = datagen.batch_sets(nsets, seed=seed)[:sample]
    workers = max(1, min(16, os.cpu_count() or 1))
    O.dt_alpha_points(sets[0])   # library load and predicate set-up, not timed
    with ThreadPoolExecutor(workers) as pool:
        t0 = time.perf_counter()
        pts = sum(pool.map(O.dt_alpha_points, sets))
        dt = time.perf_counter() - t0
    return {"cpu_sets_per_s": round(len(sets) / dt, 1), "cpu_cores": workers, "cpu_sample_sets": len(sets),
            "cpu_sample": f"first {len(sets)} sets ({pts} points), oracle DT + alpha, one set per task, "
                          f"{workers} threads, {dt:.2f} s"}


def host_to_host(G, xy, T, reps=3):
    """The reference-named calls with host arrays in and out: gpudtComputeDT, then
    gpudtComputeAlpha on its output.  Median of `reps` runs after one warm-up."""
    import ctypes as C
    import numpy as np
    L = G.lib()
    L.gpudtComputeDT.argtypes = [C.POINTER(G._Params)]
    L.gpudtComputeDT.restype = C.POINTER(G._Output)
    L.gpudtReleaseDTOutput.argtypes = [C.POINTER(G._Output)]
    L.gpudtComputeAlpha.argtypes = [C.POINTER(G._Params), C.POINTER(G._Output), C.POINTER(C.POINTER(G._Alpha))]
    L.gpudtComputeAlpha.restype = C.c_int
    n = xy.shape[0]
    prm = G._Params(0, 0, 1, 1, n, xy.ctypes.data_as(C.POINTER(C.c_double)), 0, None, T)
    times = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        out = L.gpudtComputeDT(C.byref(prm))
        if not out:
            raise RuntimeError(L.gpudtGetLastError().decode())
        A = C.POINTER(G._Alpha)()
        if L.gpudtComputeAlpha(C.byref(prm), out, C.byref(A)) < 0:
            raise RuntimeError(L.gpudtGetLastError().decode())
        dt = time.perf_counter() - t0
        L.gpudtReleaseAlpha(A)
        L.gpudtReleaseDTOutput(out)
        if r:
            times.append(dt)
    med = float(np.median(times))
    return {"value": round(n / med / 1e6, 3), "unit": "Mpoints/s", "ms": round(med * 1e3, 2),
            "calls": "gpudtComputeDT + gpudtComputeAlpha, host arrays in and out (PCIe copies included; the "
                     "alpha call hashes the unchanged DT output and uses the copy still on the device)"}


def stage_gbs(st, a, n, T):
    """Algorithmic bytes of every stage by the scope table's formulas (P texels, N points, V ~ N
    vertices, F triangles, E edges, S = F + E spectrum entries, M master entries; a mesh pass
    reads and writes 36F + 32V, a flip test 32 B per listed triangle) over the measured stage
    time.  The flip stage holds every flip pass: the passes after shifting and missing-site
    insertion and those after frame removal."""
    P, F, E = T * T, st["nTris"], st["nEdges"]
    S, M, V = F + E, len(a["master_simplex"]), n
    listed = st.get("stageListed", [0, 0, 0])
    mesh_pass = 36 * F + 32 * V
    model = [
        ("pba", "S0-S2: 4P + 20N + 24P + 12P", 40 * P + 20 * n, st["ms_pba"]),
        ("recon", "S3-S4: 4P + 12F + 48F", 4 * P + 60 * F, st["ms_recon"]),
        ("shift", "S5 one pass 36F + 32V", mesh_pass, st["ms_shift"]),
        ("missing", "S5 one pass 36F + 32V", mesh_pass, st["ms_missing"]),
        ("boundary", "S5 one pass 36F + 32V", mesh_pass, st["ms_boundary"]),
        ("flip", "S6 32 B per triangle listed for a flip test (all passes)", 32 * (listed[0] + listed[1]),
         st["ms_flip"]),
        ("canon", "triangle records read once (24F) and output records written once (36F)", 60 * F, st["ms_canon"]),
    ]
    if st.get("ms_alpha_rho", 0) > 0:
        model += [
            ("alpha_rho", "S7: 24F + 16N*3 + 8(F+E) + 12E", 24 * F + 48 * n + 8 * (F + E) + 12 * E, st["ms_alpha_rho"]),
            ("alpha_order", "S8: 24S", 24 * S, st["ms_alpha_order"]),
            ("alpha_lists", "S9: 8N + 24E + 4F + 24M", 8 * n + 24 * E + 4 * F + 24 * M, st["ms_alpha_lists"]),
        ]
    else:
        model.append(("alpha", "S7-S9: 24F + 48N + 8(F+E) + 12E + 24S + 8N + 24E + 4F + 24M",
                      24 * F + 48 * n + 8 * (F + E) + 12 * E + 24 * S + 8 * n + 24 * E + 4 * F + 24 * M,
                      st["ms_alpha"]))
    out = {}
    for k, formula, b, ms in model:
        gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        out[k] = {"formula": formula, "bytes": int(b), "ms": round(ms, 3), "gb_s": round(gbs, 1),
                  "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def stress_c3(G, reps=3, seed=3):
    """The degenerate / clustered stress workload: 2,000,000 FP64 points, half drawn without
    replacement from a 1414 x 1414 lattice scaled to [0,1)^2 (many co-circular quadruples), half
    from 64 Gaussian clusters (sigma 0.005, centres uniform); seed 3; texture 8192 x 8192; DT +
    full alpha, points resident in HBM.  Median of `reps` runs after one warm-up, and the exact
    fallback rate of every predicate (exact evaluations / filtered ones), overall and per mesh
    stage."""
    import numpy as np
    from galpha2d_amd import datagen
    n, T = 2_000_000, 8192
    xy = datagen.lattice_clusters(n, seed=seed)
    ctx = G.Context(0)
    d = G.DeviceBuffer(xy)
    ctx.compute_device(d.ptr, n, T, with_alpha=True)
    G.synchronize()
    times, stage = [], {}
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.compute_device(d.ptr, n, T, with_alpha=True)
        G.synchronize()
        times.append(time.perf_counter() - t0)
        for k, v in ctx.stats().items():
            if k.startswith("ms_"):
                stage.setdefault(k, []).append(v)
    st = ctx.stats()
    ctx.close()
    d.free()
    med = float(np.median(times))

    def rate(exact, tests, per="tests"):
        return {"exact": int(exact), per: int(tests), "rate": round(exact / tests, 8) if tests else None}
    out = {"workload": f"{n} points: half from a 1414 x 1414 lattice (without replacement), half from 64 "
                       f"Gaussian clusters (sigma 0.005); seed {seed}; texture {T}x{T}; DT + full alpha",
           "ms": round(med * 1e3, 2), "mpoints_per_s": round(n / med / 1e6, 3), "runs": reps,
           "stage_ms": {k: round(float(np.median(v)), 3) for k, v in stage.items()},
           "triangles": st["nTris"], "frame_bulk": st["frameBulk"],
           "work": {k: st[k] for k in ("shiftBad", "deleted", "missing", "inserted", "duplicates", "insertRounds",
                                        "flipPasses", "flips", "flipListed", "earRounds", "alphaClusters",
                                        "alphaMaxCluster")},
           "fallback": {
               "incircle": rate(st["incircleExact"], st["incircleTests"]),
               "orient_exact": int(st["orientExact"]),
               "alpha_compare": rate(st["alphaExactCompares"], st["nTris"] + st["nEdges"], "entries"),
               "attachment": rate(st["alphaExactAttach"], st["nEdges"], "edges")}}
    if "stageIncircleTests" in st:
        names = ("shift, insertion and flips", "frame removal and final flips", "constraints and canonical output")
        out["fallback"]["per_stage"] = {
            nm: {"incircle": rate(st["stageIncircleExact"][i], st["stageIncircleTests"][i]),
                 "orient_exact": int(st["stageOrientExact"][i])} for i, nm in enumerate(names)}
    return out


def batched_vs_plain(G, nsets, plain_sample, seed=4):
    """Batched workload: nsets point sets, sizes log-uniform in [256, 65536], each uniform or an
    8-cluster Gaussian mix (50/50), autotuned texture per set; DT + full alpha for every set.
    Points are resident in HBM before timing; results stay on the device.  Batched = one
    galphaBatchDevice pass over all sets; plain = one galphaComputeDevice call per set, timed on
    the first `plain_sample` sets (sets/s is a rate, so a sample of the same distribution is
    enough)."""
    import numpy as np
    from galpha2d_amd import datagen
    sets = datagen.batch_sets(nsets, seed=seed)
    sizes = np.array([x.shape[0] for x in sets], dtype=np.int64)
    offs = np.zeros(nsets + 1, dtype=np.int32)
    offs[1:] = np.cumsum(sizes)
    d_xy = G.DeviceBuffer(np.ascontiguousarray(np.concatenate(sets)))
    sample = sets[:min(256, nsets)]
    del sets
    ctx = G.Context(0)
    stream = None   # the context's own stream
    G.synchronize()
    base = d_xy.ptr
    # texture autotuner on a sample (the first 256 sets), not timed; both modes use its choice
    t0 = time.perf_counter()
    dens = ctx.autotune(sample, reps=1)
    t_tune = time.perf_counter() - t0
    del sample
    # plain loop (warm up on the first 8 sets)
    ns = min(plain_sample, nsets)
    for k in range(min(8, ns)):
        ctx.compute_device(base + 16 * int(offs[k]), int(sizes[k]), ctx.texture_for(int(sizes[k])),
                           with_alpha=True, stream=stream)
    G.synchronize()
    t0 = time.perf_counter()
    for k in range(ns):
        ctx.compute_device(base + 16 * int(offs[k]), int(sizes[k]), ctx.texture_for(int(sizes[k])),
                           with_alpha=True, stream=stream)
    G.synchronize()
    t_plain = time.perf_counter() - t0
    # batched: one warm-up pass (allocations), then timed passes
    ctx.batch_device(base, offs, with_alpha=True, stream=stream)
    G.synchronize()
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        F = ctx.batch_device(base, offs, with_alpha=True, stream=stream)
    G.synchronize()
    t_batch = (time.perf_counter() - t0) / reps
    st = ctx.stats()
    ctx.close()
    d_xy.free()
    plain_rate = ns / t_plain
    batch_rate = nsets / t_batch
    return {"sets": nsets, "points": int(sizes.sum()), "triangles": int(F),
            "sets_per_s": round(batch_rate, 1), "ms_per_batch": round(t_batch * 1e3, 1),
            "plain_sets_per_s": round(plain_rate, 1), "plain_sample_sets": ns,
            "speedup_vs_plain": round(batch_rate / plain_rate, 2),
            "batch_stage_ms": {k: round(v, 1) for k, v in st.items() if k.startswith("ms_")},
            "autotuned_pixels_per_point": [float(x) for x in dens], "autotune_s": round(t_tune, 1),
            "workload": f"{nsets} sets, sizes log-uniform in [256, 65536], uniform or 8-cluster Gaussian "
                        f"(seed {seed}), autotuned textures, DT + alpha, device-resident"}


def serving(G, qps_list, duration_s, seed=5, slots=4, budget_gb=192, max_group=64, max_group_points=2_000_000):
    """Serving workload: Poisson arrivals at a fixed rate, sizes log-uniform in [1,000, 2,000,000]
    i.i.d. uniform points (distinct windows of one seeded pool), DT + alpha per request, results
    copied to host.  Latency is measured inside the library from submit to completion."""
    import threading
    import numpy as np
    from galpha2d_amd import datagen
    rng = np.random.default_rng(seed)
    pool = datagen.uniform(seed, 2_000_000 + 1)
    srv = G.Server(0, slots=slots, hbm_budget_bytes=budget_gb << 30, max_group=max_group,
                   max_group_points=max_group_points)
    # warm-up: one large and a few small requests (buffer growth)
    for n in (2_000_000, 300_000, 3000):
        t = srv.submit(pool[:n], with_alpha=True)
        srv.wait(t)
        srv.release(t)
    out = []
    for qps in qps_list:
        nreq = max(1, int(qps * duration_s))
        gaps = rng.exponential(1.0 / qps, nreq)
        sizes = np.floor(np.exp(rng.uniform(np.log(1000), np.log(2_000_000), nreq))).astype(np.int64)
        starts = rng.integers(0, pool.shape[0] - sizes + 1)
        tickets = [None] * nreq
        res = [None] * nreq
        ready = threading.Semaphore(0)

        def waiter():
            for i in range(nreq):
                ready.acquire()
                F, lat, q, g = srv.wait(tickets[i])
                srv.release(tickets[i])
                res[i] = (lat, q, g)
        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        t_next = t0
        for i in range(nreq):
            t_next += gaps[i]
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(pool[starts[i]:starts[i] + sizes[i]], with_alpha=True)
            ready.release()
        th.join()
        wall = time.perf_counter() - t0
        lat = np.array([r[0] for r in res])
        grp = np.array([r[2] for r in res])
        by_size = {}
        for name, lo, hi in (("<100K", 0, 100_000), ("100K-1M", 100_000, 1_000_000), (">=1M", 1_000_000, 1 << 62)):
            sel = (sizes >= lo) & (sizes < hi)
            if sel.any():
                by_size[name] = {"requests": int(sel.sum()), "p50_ms": round(float(np.percentile(lat[sel], 50)), 2),
                                 "p99_ms": round(float(np.percentile(lat[sel], 99)), 2)}
        out.append({"qps_offered": qps, "requests": nreq, "points": int(sizes.sum()),
                    "sets_per_s": round(nreq / wall, 1), "p50_ms": round(float(np.percentile(lat, 50)), 2),
                    "p99_ms": round(float(np.percentile(lat, 99)), 2),
                    "mean_queue_ms": round(float(np.mean([r[1] for r in res])), 2),
                    "mean_group": round(float(grp.mean()), 2), "by_size": by_size})
    st = srv.stats()
    srv.close()
    return {"runs": out, "slots": slots, "hbm_budget_gb": budget_gb, "max_group": max_group,
            "max_group_points": max_group_points,
            "held_gb_end": round(st["held_bytes"] / 2**30, 2),
            "workload": f"Poisson arrivals, sizes log-uniform in [1000, 2000000] uniform points (seed {seed}), "
                        f"DT + alpha, results copied to host"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_POINTS)
    ap.add_argument("--texture", type=int, default=TEXTURE)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-host-to-host", action="store_true",
                    help="skip the host-to-host runs (profiled commands count only the timed workload)")
    ap.add_argument("--probe", default="k_edge_intervals", choices=sorted(ALGO_BYTES),
                    help="kernel timed for the roofline figure (default: the hand-written kernel with "
                         "the largest share of GPU time in profiles/c2_kernel_stats.txt)")
    ap.add_argument("--batched-sets", type=int, default=4096, help="sets of the batched workload (0: skip)")
    ap.add_argument("--plain-sample", type=int, default=256, help="sets timed for the plain per-set loop")
    ap.add_argument("--serve-qps", default="50,150", help="offered request rates of the serving runs ('' skips)")
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--no-c3", action="store_true", help="skip the degenerate / clustered stress workload")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB device-to-device once, so that a counter run sees a plain "
                         "copy of known size beside the pipeline kernels")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent point sets would shard across GPUs with no collective)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import galpha2d_amd as G
    from galpha2d_amd import datagen

    n, T = args.n, args.texture
    xy = datagen.uniform(SEED, n)
    ctx = G.Context(0)
    # device memory and synchronization through the library's own HIP runtime
    d_xy = G.DeviceBuffer(xy)
    stream = None   # the context's own stream
    G.synchronize()
    if args.copy_calibration:
        src = G.DeviceBuffer(np.ones(64 * 1024 * 1024, dtype=np.int32))
        dst = G.DeviceBuffer(nbytes=src.nbytes)
        G.synchronize()
        dst.copy_from(src)
        G.synchronize()
        src.free()
        dst.free()

    def step():
        return ctx.compute_device(d_xy.ptr, n, T, with_alpha=True, stream=stream)

    for _ in range(args.warmup):
        step()
    G.synchronize()
    stage_ms = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        s = ctx.stats()
        for k, v in s.items():
            if k.startswith("ms_"):
                stage_ms[k] = stage_ms.get(k, 0.0) + v
    G.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = n * args.steps / elapsed / 1e6
    st = ctx.stats()

    # roofline of the dominant hand-written kernel (largest share of GPU time after rocPRIM's
    # sort passes in the rocprof summary of this workload): algorithmic bytes of the probed run /
    # total time of its launches in that run
    probe = ctx.kernel_probe(args.probe, lambda: step())
    pst = ctx.stats()
    gpu_alpha = ctx.alpha()
    gpu_tris = ctx.triangles() if not args.no_cpu_baseline else None
    stages = stage_gbs(pst, gpu_alpha, n, T)
    roof = None
    if probe and probe["launches"] > 0:
        model = ALGO_BYTES[args.probe]
        total_bytes = model["bytes"](pst, T, n)
        per_launch = total_bytes / probe["launches"]
        avg_ms = probe["total_ms"] / probe["launches"]
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        traffic = measured_traffic(args.probe)
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": args.probe, "launches": probe["launches"], "avg_launch_ms": round(avg_ms, 4),
                "algorithmic_bytes_per_launch": round(per_launch), "algorithmic_model": model["model"],
                "traffic_source": "profiles/c2_hbm_traffic.txt (FETCH_SIZE and WRITE_SIZE passes, scaled by "
                                  "known-size streaming kernels)"}
        if "gather_bytes" in model:
            gb = model["gather_bytes"](pst, T, n) / probe["launches"]
            roof["gather_model"] = model["gather_model"]
            roof["gather_model_achieved"] = round(gb / (avg_ms * 1e-3) / 1e9, 2)

    out = {
        "metric": "Mpoints/s DT+alpha at 1M pts; batched sets/s and p50/p99 latency vs plain",
        "value": round(value, 3), "unit": "Mpoints/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n} uniform FP64 points in [0,1)^2 (seed {SEED}), texture {T}x{T}, "
                               f"DT + full alpha spectrum", "points": n, "texture": T,
                   "triangles": st["nTris"], "spectrum": st["specSize"]},
        "stage_ms": {k: round(v / args.steps, 3) for k, v in stage_ms.items()},
        "precision": {"incircle_tests": st["incircleTests"], "incircle_exact": st["incircleExact"],
                      "incircle_exact_rate": round(st["incircleExact"] / max(1, st["incircleTests"]), 8),
                      "orient_exact": st["orientExact"], "alpha_exact_compares": st["alphaExactCompares"],
                      "alpha_exact_attach": st["alphaExactAttach"]},
        "roofline": roof,
        "stage_gbs": stages,
        "work": {k: st[k] for k in ("reconTris", "shiftBad", "deleted", "missing", "inserted", "duplicates",
                                     "insertRounds", "flipPasses", "flips", "flipListed", "frameBulk", "earRounds",
                                     "islandSites", "alphaClusters", "alphaMaxCluster")},
    }
    ctx.close()
    d_xy.free()
    if not args.no_host_to_host:
        out["host_to_host"] = host_to_host(G, xy, T)
    if args.batched_sets > 0:
        out["batched"] = batched_vs_plain(G, args.batched_sets, args.plain_sample)
        if not args.no_cpu_baseline:
            out["batched"].update(cpu_batched(args.batched_sets, min(args.batched_sets, 1024)))
    if args.serve_qps:
        out["serving"] = serving(G, [float(q) for q in args.serve_qps.split(",")], args.serve_seconds)
    if not args.no_c3:
        out["c3"] = stress_c3(G)
    if not args.no_cpu_baseline:
        base, ref_tris, ref_alpha = cpu_baseline(n, SEED)
        out["cpu_baseline"] = base
        out["oracle_check"] = oracle_check(gpu_tris, gpu_alpha, ref_tris, ref_alpha)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
