"""Front-end benchmark: intersecting edge-face pairs per second on one MI355X.

Workload: two icospheres at subdivision 7 (327,680 triangles each), every
vertex scaled radially by 1 + U(-0.02, 0.02) (seeds 1 and 2), the second
rotated by a random rotation (seed 3) and moved +0.5 in x; float64
coordinates, synthetic data.  A step is one full front end on the GPU
(shift, polygon-soup numbering, both LBVHs, traversal, edge-face tests, face
planes, narrow phase, sorted registry, status) with the input meshes already
resident in HBM.

Prints one JSON line.  The CPU baseline is MCUT's own front end timed on this
host on the same input (the reference compiled from its sources with its stage
timers on, a helper thread per usable core but one, and single-threaded); the
single-threaded C restatement of the front end (oracle/) is timed beside it.

The line also carries the batched mix ("batched"): 4096 seeded sphere/box
pairs of uneven size, one batched dispatch against one plain dispatch per
pair, both over inputs resident in HBM; and the served stream ("served"): a
resident source mesh cut by crack-sheet requests arriving as a Poisson
process, p50/p99 latency and throughput at each rate; and the precision stage
("precision"): the degenerate-heavy heightfield sectioning and the predicate
stage alone with its per-stage decision rates.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


# stage -> the kernel that carries it (stages of one kernel are timed exactly
# by the events around that kernel on its stream).  The coordinate sums are
# five kernels (run items, chain, verification, serial fallback).
STAGE_KERNEL = {
    "traversal": "k_traverse_wave",
    "tests": "k_tests",
    "pair_filter": "k_filter_pairs",
    "face_planes": "k_planes + k_big_polygons",
    "narrow": "k_narrow + k_narrow_exact",
}


def stage_bytes(name, n):
    """Algorithmic HBM bytes of one front-end stage: every datum it must read
    or write, moved once."""
    V, F, K, E = n["V"], n["F"], n["K"], n["E"]
    C, CR, T, I, FC = n["pairs"], n["raw_pairs"], n["tests"], n["registry"], n["cand"]
    NODES = n["nodes"]
    table = {
        # every coordinate read once
        "coordinate_sums": 24 * V,
        "shift_params": 0,
        # coordinates read and shifted coordinates written; face lists read,
        # shifted face boxes written
        "shift_boxes": 24 * V + 24 * V + 4 * K + 48 * F,
        # face lists read, corner -> edge written, edge table written
        "topology": 4 * K + 4 * K + 16 * E,
        # face lists and coordinates read; Morton keys, face ids and raw boxes
        # written; four 8-bit radix passes over (key, face)
        "bvh_sort": 4 * K + 24 * V + 56 * F + 4 * 16 * F,
        # leaf boxes and leaf order read, node boxes written
        "bvh_refit": 52 * F + 48 * NODES,
        # source leaves (float box and face id) and the cut tree (float node
        # boxes, leaf face ids) read once, raw pairs written
        "traversal": 28 * n["Fs"] + 24 * n["nodes_cut"] + 4 * (F - n["Fs"]) + 8 * CR,
        # raw pairs and both shifted boxes read; pairs, pair hash and
        # candidate faces written
        "pair_filter": CR * (8 + 96) + 16 * C + 4 * FC,
        # candidate faces: face list and vertices read, plane written
        "face_planes": FC * (72 + 40),
        # pairs read, tests written, per test: edge endpoints and face box
        "tests": 8 * C + 8 * T + T * (16 + 48 + 48),
        # per test: (edge, face), endpoint ids and coordinates, plane, face
        # vertices; registry entries written (key 8 B + point 24 B)
        "narrow": T * (8 + 8 + 48 + 40 + 12 + 72) + 32 * I,
        # registry keys sorted once, entries gathered, status keys
        "registry": 32 * I + 40 * I + 16 * I,
    }
    return table.get(name, 0)


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(src, cut, registry):
    """MCUT's own front end on this host, on the same pair: the reference
    compiled from its sources with its stage timers on (oracle/ref_timing.py),
    one context with a helper thread per usable core but one
    (mcCreateContextWithHelpers, the pool is capped at the hardware
    concurrency, tpool.h:284-288), 1 warm-up dispatch, median of 5; and the
    same single-threaded (plain mcCreateContext: no helpers), median of 3.
    The single-threaded C restatement (oracle/) is timed beside them."""
    from oracle import ref_timing as rt
    from oracle import oracle as om
    out = {}
    cores = rt.usable_cores()
    if rt.available():
        helpers = max(1, cores - 1)
        workdir = "/dev/shm" if os.path.isdir("/dev/shm") else None  # MCUT's debug meshes land in RAM
        mt = rt.time_frontend(src, cut, helpers=helpers, runs=5, warmup=1, workdir=workdir)
        progress(f"cpu baseline: reference {mt['median_ms']:.0f} ms with {helpers} helpers")
        st = rt.time_frontend(src, cut, helpers=0, runs=3, warmup=1, workdir=workdir)
        progress(f"cpu baseline: reference {st['median_ms']:.0f} ms single-threaded")
        out = {"value": registry / (mt["median_ms"] * 1e-3), "unit": "pairs/s", "cores": cores, "kind": "reference",
               "sample": f"MCUT's front end (sum of its own stage timers, preproc.cpp:2304 -> kernel.cpp:3229) on the "
                         f"whole pair, mcCreateContextWithHelpers with {helpers} helpers, median of 5 after 1 warm-up "
                         f"({mt['median_ms']:.0f} ms; min {mt['min_ms']}, max {mt['max_ms']})",
               "cpu_model": rt.cpu_model(), "nproc": os.cpu_count(), "helpers": helpers,
               "frontend_ms": mt["median_ms"], "runs_ms": mt["runs_ms"], "stage_ms": mt["stage_ms"],
               "single_thread": {"value": registry / (st["median_ms"] * 1e-3), "frontend_ms": st["median_ms"],
                                 "runs_ms": st["runs_ms"], "stage_ms": st["stage_ms"]}}
    t0 = time.perf_counter()
    O = om.frontend(src, cut, detail=False)
    dt = time.perf_counter() - t0
    assert O["counts"]["registry"] == registry, (O["counts"], registry)
    port = {"value": registry / dt, "unit": "pairs/s", "cores": 1, "kind": "port",
            "sample": f"the C restatement of the front end once on the whole pair ({dt:.2f} s, single thread)"}
    if not out:
        return port
    out["port"] = port
    return out


def bench_batch(args, torch, frontend, mg):
    """Batched mix: num_pairs seeded (displaced geodesic sphere, rotated
    tessellated box) pairs resident in HBM.  One mcgpu_dispatch_batch call
    against a loop of one mcgpu_dispatch_frontend call per pair over the same
    device buffers; per-pair statuses and registry counts of the two must
    agree."""
    t0 = time.perf_counter()
    pairs = mg.config4_pairs(args.batch_pairs, seed=5)
    P = frontend.pack_batch(pairs)
    gen_s = time.perf_counter() - t0
    progress(f"batched mix: {args.batch_pairs} pairs generated in {gen_s:.1f} s")
    del pairs
    dev = torch.device("cuda:0")
    Vs, Fs_, _ = P["src"]
    Vc, Fc_, _ = P["cut"]
    tVs = torch.from_numpy(Vs).to(dev)
    tFs = torch.from_numpy(Fs_.view(np.int32)).to(dev)
    tVc = torch.from_numpy(Vc).to(dev)
    tFc = torch.from_numpy(Fc_.view(np.int32)).to(dev)
    torch.cuda.synchronize()
    nfs, nfc = int(P["src_fo"][-1]), int(P["cut_fo"][-1])
    ms = frontend.device_mesh(tVs.data_ptr(), tFs.data_ptr(), None, Vs.shape[0], nfs)
    mc = frontend.device_mesh(tVc.data_ptr(), tFc.data_ptr(), None, Vc.shape[0], nfc)
    offs = [P["src_vo"], P["src_fo"], P["cut_vo"], P["cut_fo"]]
    np_ = P["num_pairs"]
    bctx = frontend.Context(0)
    if args.batch_split:
        bctx.set_batch_split(args.batch_split)
    B = bctx.dispatch_batch_device(np_, ms, mc, offs)  # warm-up: sizes the buffers
    torch.cuda.synchronize()
    progress(f"batched mix: warm-up batch done ({B.counts})")
    tuning = None
    if args.shape:
        bctx.set_tuning(*args.shape)
        tuning = dict(wave_leaves=args.shape[0], tree_stride=args.shape[1], queues=args.shape[2],
                      chunk=args.shape[3], forced=True)
    elif not args.no_autotune:  # the traversal shape for this batch class, before the timed batches
        tuning = bctx.autotune()
        progress(f"batched mix: autotuned {tuning}")
    bt = []
    for _ in range(args.batch_steps):
        t1 = time.perf_counter()
        B = bctx.dispatch_batch_device(np_, ms, mc, offs)
        bt.append(time.perf_counter() - t1)
    batch_s = min(bt)
    bctx.set_stage_timing(True)  # stage events only outside the timed batches
    bstages = bctx.dispatch_batch_device(np_, ms, mc, offs).stage_ms()
    bctx.set_stage_timing(False)
    progress(f"batched mix: batch {batch_s * 1e3:.1f} ms")
    B = bctx.dispatch_batch_device(np_, ms, mc, offs)  # the per-pair outputs, read before the next dispatch
    reg_b = np.diff(B.reg_offsets.astype(np.int64))
    status_b = B.status.copy()
    e_b, g_b, p_b = (x.copy() for x in B.registry())
    # the same batch replayed from HIP graphs
    bctx.set_graphs(True)
    G = bctx.dispatch_batch_device(np_, ms, mc, offs)  # records the program, then replays it
    gt = []
    for _ in range(args.batch_steps):
        t1 = time.perf_counter()
        G = bctx.dispatch_batch_device(np_, ms, mc, offs)
        gt.append(time.perf_counter() - t1)
    ge, gg, gp = G.registry()
    graph_b = {"batch_ms": min(gt) * 1e3, "device_ms": G.device_ms, **bctx.graph_stats(),
               "same_as_eager": bool(np.array_equal(G.status, status_b) and np.array_equal(ge, e_b)
                                     and np.array_equal(gg, g_b) and np.array_equal(gp.view(np.uint64),
                                                                                    p_b.view(np.uint64)))}
    progress(f"batched mix: graph replay {graph_b['batch_ms']:.1f} ms")
    counts_b = dict(B.counts)
    device_ms_b = B.device_ms
    bctx.close()

    I = int(reg_b.sum())
    out = {"workload": f"{np_} pairs: geodesic sphere (logU 200-200k tris, +-2% radial noise) vs rotated "
                       f"tessellated box (logU 12-12k tris), seed 5",
           "faces": nfs + nfc, "registry": I, "pairs_not_success": int(np.count_nonzero(status_b)),
           "batch_ms": batch_s * 1e3, "batch_device_ms": device_ms_b, "batch_pairs_per_s": I / batch_s,
           "counts": counts_b, "stage_ms": bstages, "autotune": tuning, "graph_replay": graph_b, "generate_s": gen_s}
    if args.no_plain_loop:  # profiling runs: the batch alone
        return out
    pctx = frontend.Context(0)

    def plain(p):
        vs0, vs1, fs0, fs1 = int(P["src_vo"][p]), int(P["src_vo"][p + 1]), int(P["src_fo"][p]), int(P["src_fo"][p + 1])
        vc0, vc1, fc0, fc1 = int(P["cut_vo"][p]), int(P["cut_vo"][p + 1]), int(P["cut_fo"][p]), int(P["cut_fo"][p + 1])
        a = frontend.device_mesh(tVs.data_ptr() + 24 * vs0, tFs.data_ptr() + 12 * fs0, None, vs1 - vs0, fs1 - fs0)
        b = frontend.device_mesh(tVc.data_ptr() + 24 * vc0, tFc.data_ptr() + 12 * fc0, None, vc1 - vc0, fc1 - fc0)
        return pctx.dispatch_device(a, b)
    for p in range(np_):  # warm-up pass: sizes the buffers for every pair
        plain(p)
    torch.cuda.synchronize()
    progress(f"batched mix: plain warm-up pass done")
    reg_p = np.zeros(np_, np.int64)
    status_p = np.zeros(np_, np.int32)
    t1 = time.perf_counter()
    for p in range(np_):
        R = plain(p)
        reg_p[p] = R.counts["registry"]
        status_p[p] = R.status
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t1
    pctx.close()
    assert np.array_equal(reg_b, reg_p) and np.array_equal(status_b, status_p), "batch and plain results differ"
    out.update({"plain_loop_ms": plain_s * 1e3, "batched_vs_plain": plain_s / batch_s, "plain_pairs_per_s": I / plain_s})
    return out


def bench_serve(args, frontend, mg):
    """Served stream: one resident displaced geodesic sphere (n = 91, 165,620
    triangles) cut by crack-sheet requests (2k-50k triangles, a pool of 256
    seeded sheets cycled) arriving as a Poisson process, swept over rates.
    Latency = completion - arrival on the host clock; every repeat of a pool
    sheet must report the same registry count."""
    src = mg.config5_source(91)
    pool = [mg.crack_sheet(k, seed=6) for k in range(args.serve_pool)]
    mv = max(c[0].shape[0] for c in pool)
    mf = max(c[1].size // 3 for c in pool)
    srv = frontend.Server(src, mv, mf, slots=args.serve_slots, hbm_budget=args.serve_budget_gb << 30)
    progress(f"serving: {srv.slots} slots of {srv.slot_bytes / 2**20:.0f} MiB, pool of {len(pool)} sheets")
    # service time of each pool sheet alone (one in flight)
    res = srv.run(pool, np.arange(len(pool)) * 0.0 + np.arange(len(pool)) * 1e-9, keep_registries=False)
    ref_counts = res["num_registry"].astype(np.int64)
    rng = np.random.default_rng(6)
    rates = []
    for lam in args.serve_rates:
        n = int(min(args.serve_max, max(50, lam * args.serve_seconds)))
        arr = np.cumsum(rng.exponential(1.0 / lam, n))
        cuts = [pool[i % len(pool)] for i in range(n)]
        r = srv.run(cuts, arr)
        lat = (r["done_s"] - r["arrival_s"]) * 1e3
        span = r["done_s"].max() - r["arrival_s"].min()
        counts = r["num_registry"].astype(np.int64)
        assert np.all(r["status"] >= 0) and np.array_equal(counts, ref_counts[np.arange(n) % len(pool)]), \
            "a request's result changed in the mix"
        rates.append({"rate_rps": lam, "requests": n, "p50_ms": float(np.percentile(lat, 50)),
                      "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                      "throughput_rps": n / span, "pairs_per_s": float(counts.sum()) / span,
                      "queue_wait_p50_ms": float(np.percentile((r["start_s"] - r["arrival_s"]) * 1e3, 50))})
        progress(f"serving: {lam} req/s -> p50 {rates[-1]['p50_ms']:.2f} ms, p99 {rates[-1]['p99_ms']:.2f} ms")
    srv.close()
    return {"workload": "resident geodesic sphere n=91 (165,620 tris, +-2% radial noise) cut by open crack sheets "
                        "(logU 2k-50k tris, side 3, random orientation, offset U(-0.6,0.6), 0.03 fBm), "
                        f"Poisson arrivals, seed 6, {len(pool)} distinct sheets cycled",
            "slots": srv.slots, "slot_mib": srv.slot_bytes / 2**20, "hbm_budget_gib": args.serve_budget_gb,
            "alone_registry_mean": float(ref_counts.mean()), "rates": rates}


def bench_precision(args, frontend, mg):
    """Degenerate-heavy sectioning: the 1024 x 1024 heightfield cut at
    z = 410/4096 (status as the reference reports it), its x/y-scaled variant
    (faces 1 x 1, valid planes, many vertices exactly on the cut plane), and
    the predicate stage alone over both: 2 T orient3d signs through the FP32
    pre-filter, the FP64 filter and exact arithmetic."""
    ctx = frontend.Context(0)
    out = {"bounds": {"f32": "16 * 2^-24 * permanent (9.54e-7 * permanent), coordinate differences in [2^-40, 2^40]",
                      "f64": "(7 + 56 eps) eps * permanent, eps = 2^-53 (7.77e-16 * permanent; shewchuk.c:424)"}}
    for name, scale in (("heightfield", 1.0), ("heightfield_scaled", 1024.0)):
        src, cut = mg.config3(scale=scale)
        R = ctx.dispatch(src, cut)
        dev = []
        for _ in range(3):
            R = ctx.dispatch(src, cut)
            dev.append(R.device_ms)
        c = dict(R.counts)
        T = c["tests"]
        P = R.predicate_stage()
        P2 = R.predicate_stage()
        P["device_ms"] = min(P["device_ms"], P2["device_ms"])
        out[name] = {"workload": f"1024 x 1024 heightfield (fBm heights in [0, 0.2] on a 2^-12 grid, seed 4), "
                                 f"x/y scale {scale:g}, cut by a 64 x 64 quad plane at z = 410/4096",
                     "status": R.status_name, "frontend_device_ms": min(dev), "counts": c,
                     "narrow_stage_rates": {"f32": c["decided_f32"] / max(1, T), "f64": c["decided_f64"] / max(1, T),
                                            "exact": c["escalated"] / max(1, T)},
                     "predicate_stage": {**P, "predicates": 2 * T,
                                         "predicates_per_s": 2 * T / (P["device_ms"] * 1e-3) if P["device_ms"] else None,
                                         "escalation_rate": P["exact"] / max(1, 2 * T),
                                         "f32_rate": P["f32"] / max(1, 2 * T)}}
        progress(f"precision: {name} {R.status_name}, {T} tests, predicate stage {P['device_ms']:.3f} ms")
    ctx.close()
    return out


def bench_mcdispatch(args, frontend, mg):
    """mcDispatch end to end (input meshes to connected components ready):
    MCUT built from its own sources against the same MCUT with its whole
    intersection front end on the GPU (mcut_amd/seam: the bvh.cpp broad phase
    and dispatch()'s front-end block over libmcgpu; its components are
    identical to MCUT's), both with MCUT's debug dumps switched off, on
    displaced icosphere pairs MCUT completes (+-0.5% radial noise, rotated,
    offset +0.5 x): median of 3 dispatches after one, with a helper thread per
    usable core but one and single-threaded.  The GPU front end alone is
    timed beside them."""
    from mcut_amd.seam import mcut_api
    from oracle import ref_timing as rt
    quiet = os.path.join(ROOT, "mcut_amd", "seam", "_build", "libmcut_quiet.so")
    gpu_lib = os.path.join(ROOT, "mcut_amd", "seam", "_build", "libmcut_gpu.so")
    if not (os.path.exists(quiet) and os.path.exists(gpu_lib)):
        return {"skipped": "MCUT builds not present"}
    helpers = max(1, rt.usable_cores() - 1)
    ref, gpu = mcut_api.Mcut(quiet), mcut_api.Mcut(gpu_lib)
    out = {"helpers": helpers, "dumps": "off (preproc.cpp:2371 switched in both builds)", "pairs": []}
    for level in args.mcd_levels:
        V, F, _ = mg.icosphere(level)
        A = mg.radial_displace(V, 0.005, seed=1)
        B = mg.radial_displace(V, 0.005, seed=2) @ mg.random_rotation(seed=3).T
        B[:, 0] += 0.5
        src, cut = (A, F, None), (np.ascontiguousarray(B), F.copy(), None)
        r = args.mcd_repeats
        row = {"workload": f"icosphere-{level} pair ({F.size // 3} tris each, +-0.5% radial noise, rotated, "
                           f"offset +0.5x), MC_DISPATCH_FILTER_ALL"}
        med = lambda t: float(np.median(t[1:])) if len(t) > 1 else float(t[0])
        row["reference_cpu_ms"] = med(ref.time_dispatch(src, cut, helpers=0, repeats=r + 1))
        row["reference_cpu_helpers_ms"] = med(ref.time_dispatch(src, cut, helpers=helpers, repeats=r + 1))
        row["gpu_frontend_mcut_ms"] = med(gpu.time_dispatch(src, cut, helpers=0, repeats=r + 1))
        row["gpu_frontend_mcut_helpers_ms"] = med(gpu.time_dispatch(src, cut, helpers=helpers, repeats=r + 1))
        ctx = frontend.Context(0)
        ctx.dispatch(src, cut)
        row["gpu_frontend_device_ms"] = min(ctx.dispatch(src, cut).device_ms for _ in range(5))
        ctx.close()
        row["speedup_helpers"] = row["reference_cpu_helpers_ms"] / row["gpu_frontend_mcut_helpers_ms"]
        row["speedup_single_thread"] = row["reference_cpu_ms"] / row["gpu_frontend_mcut_ms"]
        progress(f"mcDispatch icosphere-{level}: MCUT {row['reference_cpu_helpers_ms']:.0f} ms ({helpers} helpers), "
                 f"{row['reference_cpu_ms']:.0f} ms (1 thread); GPU front end {row['gpu_frontend_mcut_helpers_ms']:.0f} / "
                 f"{row['gpu_frontend_mcut_ms']:.0f} ms")
        out["pairs"].append(row)
    return out


def bench_single(args, torch, frontend, mg):
    """The headline: K front ends of the icosphere-7 pair, inputs resident in HBM."""
    src, cut = mg.config2(level=args.level)
    dev = torch.device("cuda:0")
    keep = []

    def to_dev(m):
        V, F, _ = m
        tv = torch.from_numpy(np.ascontiguousarray(V)).to(dev)
        tf = torch.from_numpy(np.ascontiguousarray(F).view(np.int32)).to(dev)
        keep.extend([tv, tf])
        return frontend.device_mesh(tv.data_ptr(), tf.data_ptr(), None, V.shape[0], F.size // 3)

    ms, mc = to_dev(src), to_dev(cut)
    torch.cuda.synchronize()
    progress("single pair: inputs resident")
    if args.copy_calibration:
        a = torch.ones(32 << 20, dtype=torch.float64, device=dev)  # 256 MiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    ctx = frontend.Context(0)
    for _ in range(args.warmup):
        R = ctx.dispatch_device(ms, mc)
    torch.cuda.synchronize()
    # The timed steps run the library's default traversal shape (32 source
    # leaves a block, four tree levels a step, each XCD's blocks in runs of 16),
    # so the kernel profiled under profiles/ is the one timed; the
    # autotuner's own choice is reported beside it, from a context of its own.
    tuning = None
    if args.shape:  # another fixed shape
        ctx.set_tuning(*args.shape)
        R = ctx.dispatch_device(ms, mc)
    elif not args.no_autotune:
        tctx = frontend.Context(0)
        tctx.dispatch_device(ms, mc)
        tuning = tctx.autotune()
        tctx.close()
        progress(f"single pair: autotuner's choice {tuning}")
    shape_used = ctx.tuning()
    lat = []
    # the timed steps record the traversal kernel's own dispatch timestamps
    # (two events; the roofline kernel time is their mean over these steps)
    ctx.set_stage_timing("kernel")
    trav_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        R = ctx.dispatch_device(ms, mc)
        lat.append((time.perf_counter() - s0) * 1e3)
        trav_ms.append(R.stage_ms()["traversal_kernel"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.set_stage_timing(False)
    ms_step = elapsed * 1e3 / args.steps
    I = R.counts["registry"]
    assert R.status == 0, R.status_name
    device_ms = R.device_ms

    # the same dispatches replayed from HIP graphs (mcgpu_set_graphs): the
    # shape is recorded once, then every step is a replay of its program
    gctx = frontend.Context(0)
    if args.shape:
        gctx.set_tuning(*args.shape)
    gctx.set_graphs(True)
    for _ in range(max(2, args.warmup)):
        G = gctx.dispatch_device(ms, mc)
    torch.cuda.synchronize()
    glat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        G = gctx.dispatch_device(ms, mc)
        glat.append((time.perf_counter() - s0) * 1e3)
    torch.cuda.synchronize()
    g_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    ge, gg, gp = G.registry()
    re_, rg_, rp_ = R.registry()
    graph_line = {"ms_per_step": g_ms, "device_ms": G.device_ms, "p50_ms": sorted(glat)[len(glat) // 2],
                  "value": G.counts["registry"] / (g_ms * 1e-3), **gctx.graph_stats(),
                  "same_as_eager": bool(G.status == R.status and np.array_equal(ge, re_) and np.array_equal(gg, rg_)
                                        and np.array_equal(gp.view(np.uint64), rp_.view(np.uint64)))}
    gctx.close()
    progress(f"single pair: graph replay {g_ms:.3f} ms per step ({graph_line['graph_launches']} graphs)")

    # what a drop-in pays beyond the resident-input steps: host arrays in
    # (copied over PCIe each dispatch) and the registry back to the host
    hctx = frontend.Context(0)
    if args.shape:
        hctx.set_tuning(*args.shape)
    for _ in range(2):
        H = hctx.dispatch(src, cut)
        H.registry()
    ht = []
    for _ in range(args.steps):
        h0 = time.perf_counter()
        H = hctx.dispatch(src, cut)
        he, hg, hp = H.registry()
        ht.append((time.perf_counter() - h0) * 1e3)
    hctx.close()
    h_ms = sorted(ht)[len(ht) // 2]
    host_line = {"ms_per_step_p50": h_ms, "value": H.counts["registry"] / (h_ms * 1e-3),
                 "input_bytes": int(sum(np.asarray(m[0]).nbytes + np.asarray(m[1]).nbytes for m in (src, cut))),
                 "registry_bytes": int(he.nbytes + hg.nbytes + hp.nbytes)}
    progress(f"single pair: host arrays in, registry out: {h_ms:.3f} ms per step")

    # the optional single-component check (MCGPU_DISPATCH_CHECK_COMPONENTS):
    # device time of a dispatch that also counts each mesh's components
    cc_ms = []
    for _ in range(3):
        cc_ms.append(ctx.dispatch_device(ms, mc, flags=frontend.CHECK_COMPONENTS).device_ms)
    R = ctx.dispatch_device(ms, mc)  # the timed dispatches' outputs again, for the checks below

    # per-stage device times (HIP events on the library's streams), gathered
    # from further dispatches outside the timed loop: the timed dispatches
    # record no stage events (each costs time)
    stage_tot = {}
    ctx.set_stage_timing(True)
    for _ in range(args.steps):
        Rs = ctx.dispatch_device(ms, mc)
        for k, v in Rs.stage_ms().items():
            stage_tot[k] = stage_tot.get(k, 0.0) + v
    ctx.set_stage_timing(False)
    stages = {k: v / args.steps for k, v in stage_tot.items()}
    counts = dict(R.counts)
    Fs, Fc = src[1].size // 3, cut[1].size // 3
    nodes = sum((2 << max(0, int(np.ceil(np.log2(f))))) - 1 for f in (Fs, Fc))
    n = dict(V=src[0].shape[0] + cut[0].shape[0], F=Fs + Fc, K=src[1].size + cut[1].size, E=counts["edges"],
             pairs=counts["pairs"], raw_pairs=counts["raw_pairs"], tests=counts["tests"], registry=I,
             cand=counts["cand"], nodes=nodes, Fs=Fs, nodes_cut=(2 << max(0, int(np.ceil(np.log2(Fc))))) - 1)
    # the roofline kernel: the longest stage that one kernel carries alone,
    # timed by the kernel's own dispatch timestamps where the library records
    # them (the traversal), else by the events around it on its stream
    dom = max((k for k in stages if k in STAGE_KERNEL), key=stages.get)
    dom_bytes = stage_bytes(dom, n)
    # the traversal's time: its dispatch timestamps in the timed steps
    kernel_ms = float(np.mean(trav_ms)) if dom == "traversal" else 0.0
    if kernel_ms <= 0.0:
        kernel_ms = stages[dom]
    achieved = dom_bytes / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
    lat_sorted = sorted(lat)

    line = {
        "metric": "intersecting edge-face pairs/sec (front end, one mcDispatch pair)",
        "value": I / (ms_step * 1e-3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded displaced icospheres)",
        "config": {"workload": f"icosphere-{args.level} vs rotated icosphere-{args.level} "
                               f"({n['F'] // 2} tris each, +-2% radial noise, offset +0.5x)",
                   "parallelism": "single GPU"},
        "frontend_device_ms": device_ms,
        "latency_ms": {"p10": float(np.percentile(lat, 10)), "p50": lat_sorted[len(lat) // 2],
                       "p90": float(np.percentile(lat, 90)), "p99": lat_sorted[min(len(lat) - 1, int(0.99 * len(lat)))]},
        "counts": counts,
        "stage_ms": stages,
        "traversal_shape": {k: shape_used[k] for k in ("wave_leaves", "tree_stride", "queues", "chunk")},
        "autotune": tuning,
        "graph_replay": graph_line,
        "component_check_device_ms": min(cc_ms),
        "host_arrays_and_readback": host_line,
        "roofline": {"bound": "hbm", "kernel": STAGE_KERNEL[dom], "stage": dom, "kernel_ms": kernel_ms,
                     "stage_event_ms": stages[dom],
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "alg_bytes": dom_bytes,
                     "frontend_alg_bytes": sum(stage_bytes(k, n) for k in stages)},
    }
    progress(f"single pair: {ms_step:.3f} ms per step")
    # every registry point of the GPU's own result against the exact segment /
    # plane intersection (integer-scaled exact arithmetic)
    from mcut_amd import exact
    e_, g_, p_ = R.registry()
    v0, v1, _, _ = R.edges()
    err, straddle = exact.registry_point_errors(R.ps_vertices(), exact.triangle_soup_faces(src[1], cut[1], src[0].shape[0]),
                                                v0, v1, e_, g_, p_)
    line["point_error"] = {"max_rel": float(err.max()), "points": int(err.size), "all_straddle": bool(straddle.all()),
                           "bound": 1e-12, "method": "exact rational (integer-scaled) segment/plane intersection"}
    progress(f"single pair: max point error {err.max():.3e} over {err.size} points")
    traffic_file = os.path.join(ROOT, "profiles", "roofline_traffic.json")
    if os.path.exists(traffic_file):  # HBM bytes per launch of the roofline kernel, from rocprofv3 --pmc runs
        t = json.load(open(traffic_file))
        if (t.get("kernel") == STAGE_KERNEL[dom] and t.get("workload") == line["config"]["workload"]
                and t.get("shape") == line["traversal_shape"]):
            line["roofline"]["traffic"] = t["bytes_per_launch"]
            line["roofline"]["traffic_source"] = t["source"]
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(src, cut, I)
        progress("single pair: cpu baseline done")
        line["speedup_vs_cpu_baseline"] = line["value"] / line["cpu_baseline"]["value"]
        if "single_thread" in line["cpu_baseline"]:
            line["speedup_vs_cpu_single_thread"] = line["value"] / line["cpu_baseline"]["single_thread"]["value"]
    ctx.close()
    del keep
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--level", type=int, default=7)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-pairs", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--batch-split", type=int, default=0,
                    help="percent of the CUs beside the sums' for the batch's broad-phase stream (0: default)")
    ap.add_argument("--no-plain-loop", action="store_true",
                    help="skip the per-pair plain dispatches of the batched mix (profiling runs)")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the single-pair measurement (profiling runs of the other workloads)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-mcdispatch", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the default traversal shape instead of the autotuned one")
    ap.add_argument("--shape", type=int, nargs=4, metavar=("LEAVES", "STRIDE", "QUEUES", "CHUNK"),
                    help="time this traversal shape (source leaves per block, cut levels per step, 0 (Morton order), "
                         "1 (XCD-contiguous) or 8 work queues, blocks per pop) instead of tuning one")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB device-to-device copy (three times): the known-size kernel that "
                         "scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--mcd-levels", type=int, nargs="+", default=[5, 6])
    ap.add_argument("--mcd-repeats", type=int, default=3)
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[10, 30, 100, 300, 1000, 3000])
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-max", type=int, default=1000)
    ap.add_argument("--serve-pool", type=int, default=256)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-budget-gb", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from mcut_amd import frontend, meshgen as mg

    if args.no_single:  # profiling runs of the other workloads: no headline
        line = {"metric": "intersecting edge-face pairs/sec (front end, one mcDispatch pair)", "value": None,
                "unit": "pairs/s", "n_gpus": 1, "skipped": "single pair (--no-single)"}
    else:
        line = bench_single(args, torch, frontend, mg)
    if not args.no_batch:
        line["batched"] = bench_batch(args, torch, frontend, mg)
    if not args.no_serve:
        line["served"] = bench_serve(args, frontend, mg)
    if not args.no_precision:
        line["precision"] = bench_precision(args, frontend, mg)
    if not args.no_mcdispatch:
        line["mcdispatch"] = bench_mcdispatch(args, frontend, mg)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
