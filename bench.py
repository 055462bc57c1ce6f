#!/usr/bin/env python3
"""ISM throughput on one MI355X: isovalue-pairs/s of the isosurface similarity
map of a 600x248x248 float32 volume at 100 isovalues (distance fields on the
150x62x62 grid, 256-bin joint histograms), batched form vs the plain
per-isovalue form, with the CPU restatement timed beside it.

Prints one JSON line.  Inputs are generated on the host, uploaded once and
resident in HBM before the timed region; every timed step recomputes the whole
map (extraction, distance fields, histograms, NMI).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "isosurface-analysis_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "ISM isovalue-pairs/sec + p50/p99 request latency, batched/served vs plain"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak
# FP32 VALU lane-operations per second at that peak (an FMA counts 2 FLOP)
VALU_LANES_PER_S = FP32_PEAK_TFLOPS * 1e12 / 2
# HBM bytes per launch of the histogram kernel and of the LBVH walk on the
# default workload, from two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE)
# each scaled by what the counter reads for the plain copy kernel over 1 GiB
# (FETCH_SIZE x 1.9998, WRITE_SIZE x 1.0); commands and per-kernel table in
# profiles/ism_c2_600x248x248_k100_hbm_traffic.md.  The profiler cannot run
# inside this process, so the measured figures are carried here.
HIST_TRAFFIC_DEFAULT = 13596867346
WALK_TRAFFIC_DEFAULT = 5131639219


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--mode", choices=["compat", "fixed"], default="compat")
    p.add_argument("--plain-steps", type=int, default=1)
    p.add_argument("--brute-isos", type=int, default=3,
                   help="isovalues whose LBVH fields are checked against brute force")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--dims", default="600,248,248")
    p.add_argument("--k", type=int, default=100)
    p.add_argument("--ds", type=int, default=4)
    p.add_argument("--bins", type=int, default=256)
    p.add_argument("--no-serve", action="store_true")
    p.add_argument("--no-batch", action="store_true")
    p.add_argument("--copy-bytes", type=int, default=0,
                   help="run the plain copy kernel over this many bytes once before the timed "
                        "steps (reference dispatch for HBM counter calibration)")
    p.add_argument("--batch-volumes", type=int, default=48)
    p.add_argument("--batch-steps", type=int, default=2)
    p.add_argument("--serve-steps", type=int, default=24)
    p.add_argument("--serve-slots", type=int, default=4)
    p.add_argument("--serve-requests", type=int, default=1000,
                   help="requests per arrival rate (p50/p99 over at least 1000)")
    p.add_argument("--serve-rates", type=float, nargs="+", default=[50.0, 150.0, 400.0])
    return p.parse_args()


def plain_ism(ctx, isoamd, d_vol, dims, isos, ds, H, flags, vbuf, fields):
    """The drop-in form: one synchronous call chain per isovalue, as the
    reference's Single::run loop does, then one histogram call."""
    X, Y, Z = dims
    F = (X // ds) * (Y // ds) * (Z // ds)
    m = 0
    for iso in isos:
        out, n, _ = ctx.mc_extract(d_vol, dims, float(iso), flags, out=vbuf)
        if n == 0:
            continue
        dst = isoamd.DeviceArray(ctx, fields.offset(m * F), (F,), np.float32)
        ctx.df_points(out, n, dims, ds, isoamd.DF_LBVH, out=dst)
        m += 1
    smap, _, _, _ = ctx.hist_nmi(fields, m, F, H, flags=flags)
    ctx.synchronize()
    return m, smap


def cpu_baseline(img, dims, isos, ds, H, mode, cores, gpu_fields=None, gpu_map=None, ids=None):
    """The oracle (a C restatement of the reference) on a bounded sample of the
    same workload, extrapolated to the full map.  With the GPU's fields and map
    it also reports, for the sampled isovalues at full size, how many samples
    of the distance fields differ, how many voxels change joint-histogram bin
    (the oracle's binning of the GPU fields against its own) and the largest
    relative NMI difference."""
    import concurrent.futures as cf

    import oracle

    X, Y, Z = dims
    ncls = oracle.mc_ncls_compat(X, Y, Z)
    sample_isos = [isos[i] for i in np.linspace(0, len(isos) - 1, min(cores, 8)).astype(int)]
    vol = img.ravel() if mode == "compat" else np.ascontiguousarray(img.transpose(2, 1, 0)).ravel()

    def mc(iso):
        if mode == "compat":
            return oracle.mc(vol, dims, iso, ncls=ncls)[0]
        return oracle.mc(vol, dims, iso, interior=True)[0]

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(len(sample_isos)) as ex:
        surfaces = list(ex.map(mc, sample_isos))
    t_mc = (time.perf_counter() - t0) / len(sample_isos)  # per isovalue, all cores busy
    pts = oracle.normalize(surfaces[len(surfaces) // 2], dims)
    t0 = time.perf_counter()
    # the reference's search: a 1500-point pre-sample bound, then the tree
    f1 = oracle.df_tree(pts, dims, ds, threads=cores, presample=True)
    t_df = time.perf_counter() - t0
    fields = np.stack([f1, oracle.df_tree(oracle.normalize(surfaces[0], dims), dims, ds,
                                          threads=cores, presample=True)])
    reps = 4
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.ism(fields, H)  # 4 ordered pairs per call
    t_pair = (time.perf_counter() - t0) / (4 * reps)
    npairs = len(isos) * (len(isos) - 1) // 2
    total = len(isos) * (t_mc + t_df) + npairs * t_pair
    check = None
    if gpu_fields is not None:
        iso_idx = [int(np.linspace(0, len(isos) - 1, min(cores, 8)).astype(int)[len(surfaces) // 2]),
                   int(np.linspace(0, len(isos) - 1, min(cores, 8)).astype(int)[0])]
        rows = [list(ids).index(i) for i in iso_idx]
        g = np.stack([gpu_fields[r] for r in rows])
        mism = int(np.count_nonzero(g != fields))
        flips, err = 0, 0.0
        want, _ = oracle.ism(fields, H)  # the reference's float-order NMI
        for a_ in range(2):
            for b_ in range(2):
                mn_a, mx_a = oracle.interval(fields[a_])
                mn_b, mx_b = oracle.interval(fields[b_])
                h_o = oracle.hist_pair(fields[a_], fields[b_], H, mn_a, mx_a, mn_b, mx_b)
                gmn_a, gmx_a = oracle.interval(g[a_])
                gmn_b, gmx_b = oracle.interval(g[b_])
                h_g = oracle.hist_pair(g[a_], g[b_], H, gmn_a, gmx_a, gmn_b, gmx_b)
                flips += int(np.abs(h_o.astype(np.int64) - h_g.astype(np.int64)).sum() // 2)
                got = float(gpu_map[rows[a_], rows[b_]])
                ref = float(want[a_, b_])
                err = max(err, abs(got - ref) / max(abs(ref), 1e-30))
        check = {"isovalue_indices": iso_idx, "df_samples": int(fields.size),
                 "df_mismatch_samples": mism, "pairs": 4, "bin_flips": flips,
                 "nmi_max_rel_err": err}
    return {
        "value": npairs / total,
        "unit": "isovalue-pairs/s",
        "cores": cores,
        "cpu_model": _cpu_model(),
        "kind": "port",
        "sample": (f"oracle marching cubes at {len(sample_isos)} isovalues (one per core), "
                   f"oracle distance field of 1 isosurface over all {fields.shape[1]} samples "
                   f"in the reference's search order (1500-point pre-sample bound, then the "
                   f"tree), oracle histogram + reference-order NMI of 16 ordered pairs; "
                   f"extrapolated to {len(isos)} isovalues and {npairs} pairs"),
        "seconds_per": {"extract": t_mc, "df": t_df, "pair": t_pair},
        "full_size_check": check,
    }


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def nmi_vs_reference(ctx, isoamd, fields_d, gpu_fields, gpu_map, H, cores):
    """The GPU map against the reference's float-order NMI (JointHistogram.cu:
    24-60) recomputed by the oracle over the GPU's fields: every ordered pair.
    Also how far that float order itself sits from the exact value, how far a
    logf 1 ulp off (CUDA documents <= 1 ulp) could move entries, and bin flips:
    voxels binned differently by the GPU than by the oracle on sampled pairs."""
    import oracle

    m, F = gpu_fields.shape
    t0 = time.perf_counter()
    m32, m64 = oracle.ism(gpu_fields, H)
    secs = time.perf_counter() - t0
    g = gpu_map.astype(np.float64)
    rel = np.abs(g - m32) / np.maximum(np.abs(m32.astype(np.float64)), 1e-30)
    own = np.abs(m32.astype(np.float64) - m64) / np.maximum(np.abs(m64), 1e-30)
    rng = np.random.default_rng(5)
    pairs = [tuple(int(q) for q in rng.integers(0, m, 2)) for _ in range(8)]
    flips, sens = 0, 0.0
    for i, j in pairs:
        d = ctx.to_device(np.stack([gpu_fields[i], gpu_fields[j]]))
        _, joint, _, _ = ctx.hist_nmi(d, 2, F, H, want_joint=True)
        jg = joint.numpy()[0, 1].astype(np.int64)
        d.free()
        jo = oracle.hist_pair(gpu_fields[i], gpu_fields[j], H, *oracle.interval(gpu_fields[i]),
                              *oracle.interval(gpu_fields[j])).astype(np.int64)
        flips += int(np.abs(jg - jo).sum() // 2)
        v0 = oracle.nmi_ref32(jo, F)
        for seed in range(1, 5):
            sens = max(sens, abs(oracle.nmi_ref32_ulp(jo, F, seed) - v0) / max(abs(v0), 1e-30))
    return {"entries": int(m * m), "max_rel": float(rel.max()),
            "n_over_1e-5": int((rel > 1e-5).sum()), "flips": flips, "flip_pairs": len(pairs),
            "reference_order_vs_exact": {"max_rel": float(own.max()),
                                         "n_over_1e-5": int((own > 1e-5).sum())},
            "logf_1ulp_max_rel_change": sens, "oracle_seconds": round(secs, 2),
            "oracle_threads": cores}


def serve_bench(a, isoamd):
    """Serving form over the multivariate time-varying set: 3 variables x S
    steps of 256x256x128 volumes resident in HBM, a Poisson stream mixing
    ISMs (64 isovalues), TSMs (consecutive steps) and VSMs (two variables),
    cold cache per arrival rate, p50/p99 latency from arrival to result."""
    X, Y, Z = 256, 256, 128
    nvars, nsteps = 3, a.serve_steps
    t0 = time.perf_counter()
    vols = {(v, t): isoamd.gen_timeseries(X, Y, Z, v, t, nblobs=10, seed=5)
            for v in range(nvars) for t in range(nsteps)}
    log(f"serving set: {len(vols)} volumes generated in {time.perf_counter() - t0:.1f}s")
    rng = np.random.default_rng(11)
    out = []
    for rate in a.serve_rates:
        reqs = []
        for _ in range(a.serve_requests):
            u = rng.random()
            v, t = int(rng.integers(nvars)), int(rng.integers(nsteps - 1))
            if u < 0.2:
                reqs.append((isoamd.REQ_ISM, v, t, v, t))
            elif u < 0.6:
                reqs.append((isoamd.REQ_TSM, v, t, v, t + 1))
            else:
                v2 = (v + 1 + int(rng.integers(nvars - 1))) % nvars
                reqs.append((isoamd.REQ_VSM, v, t, v2, t))
        gaps = rng.exponential(1.0 / rate, len(reqs))
        with isoamd.Server(slots=a.serve_slots, k=64, ds=2, H_ism=128, H_multi=128, reps=16,
                           mem_budget=96 << 30) as s:
            for (v, t), img in vols.items():
                s.add_volume(v, t, img)
            s.start()
            t_start = time.perf_counter()
            due = t_start
            tickets = []
            for r, g in zip(reqs, gaps):
                due += g
                d = due - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets.append(s.submit(*r))
            res = [s.wait(tk) for tk in tickets]
            st = s.stats()
        lat = np.array([(i.t_done - i.t_submit) * 1e3 for _, _, i in res])
        span = max(i.t_done for _, _, i in res) - min(i.t_submit for _, _, i in res)
        pairs = sum((i.m * (i.m - 1) // 2) if r[0] == isoamd.REQ_ISM else i.m * i.m
                    for r, (_, _, i) in zip(reqs, res))
        # pairs the GPU computed (every cross map, and the maps of the ISMs
        # a request computed) and pairs served from the cache (ISM requests
        # answered without computing)
        cross = sum(i.m * i.m for r, (_, _, i) in zip(reqs, res) if r[0] != isoamd.REQ_ISM)
        ism_computed_pairs = sum(i.computed_pairs for _, _, i in res)
        cached = sum(i.m * (i.m - 1) // 2 for r, (_, _, i) in zip(reqs, res)
                     if r[0] == isoamd.REQ_ISM and i.computed == 0)
        cold = lat[[i.computed > 0 for _, _, i in res]]
        warm = lat[[i.computed == 0 for _, _, i in res]]
        by_kind = {}
        for name, kind in (("ism", isoamd.REQ_ISM), ("tsm", isoamd.REQ_TSM), ("vsm", isoamd.REQ_VSM)):
            l = lat[[r[0] == kind for r in reqs]]
            if l.size:
                by_kind[name] = {"n": int(l.size), "p50": round(float(np.percentile(l, 50)), 2),
                                 "p99": round(float(np.percentile(l, 99)), 2)}
        row = {"rate_per_s": rate, "requests": len(reqs), "seconds": round(span, 3),
               "throughput_req_s": round(len(reqs) / span, 1), "pairs_per_s": round(pairs / span, 1),
               "p50_ms": round(float(np.percentile(lat, 50)), 2),
               "p99_ms": round(float(np.percentile(lat, 99)), 2), "by_kind": by_kind,
               "computed_pairs_per_s": round((ism_computed_pairs + cross) / span, 1),
               "cache_served_pairs_per_s": round(cached / span, 1),
               "cold": {"n": int(cold.size),
                        "p50_ms": round(float(np.percentile(cold, 50)), 2) if cold.size else None,
                        "p99_ms": round(float(np.percentile(cold, 99)), 2) if cold.size else None},
               "warm": {"n": int(warm.size),
                        "p50_ms": round(float(np.percentile(warm, 50)), 2) if warm.size else None,
                        "p99_ms": round(float(np.percentile(warm, 99)), 2) if warm.size else None},
               "isms_computed": int(st.ism_computed), "cache_hits": int(st.cache_hits)}
        log(f"serve rate {rate}/s: {row}")
        out.append(row)
    return {"volumes": f"{nvars} vars x {nsteps} steps x {X}x{Y}x{Z}", "slots": a.serve_slots,
            "ism": "k=64, ds=2, 128 bins", "cross_maps": "16x16 representatives, 128 bins",
            "mix": "20% ISM, 40% TSM, 40% VSM, Poisson arrivals, cold cache per rate",
            "rates": out}


def batch_bench(a, isoamd):
    """Many volumes of uneven size in one call (iso_ism_batch): Gaussian-blob
    volumes (8 blobs, seeds 1000+v), axes uniform in 48..256, 32..128 evenly
    spaced isovalues each, ds=2, 128 bins, marching cubes (compat).  Compared
    with a loop of iso_ism_volume over the same volumes."""
    rng = np.random.default_rng(7)
    nv = a.batch_volumes
    dims_list = [tuple(int(v) for v in rng.integers(48, 257, 3)) for _ in range(nv)]
    ks = [int(v) for v in rng.integers(32, 129, nv)]
    t0 = time.perf_counter()
    ctx = isoamd.Context(0)
    vols, isos_list = [], []
    for v, (d, k) in enumerate(zip(dims_list, ks)):
        img = isoamd.gen_blobs(*d, nblobs=8, noise=0.01, seed=1000 + v)
        lo, hi = float(img.min()), float(img.max())
        isos_list.append(np.array([lo + (i + 1) * (hi - lo) / (k + 1) for i in range(k)],
                                  np.float32))
        vols.append(ctx.to_device(img))
    log(f"batch set: {nv} volumes generated in {time.perf_counter() - t0:.1f}s")
    # the pair stage runs as a captured HIP graph: captured by the warm-up
    # call, replayed by the timed ones
    run = lambda: ctx.ism_batch(vols, dims_list, isos_list, isoamd.EXTRACT_MC, 2, 128,
                                isoamd.DF_LBVH, isoamd.COMPAT_ALL | isoamd.BATCH_GRAPH)
    maps, st = run()
    ctx.synchronize()
    times = []
    for i in range(a.batch_steps):
        s0 = time.perf_counter()
        maps, st = run()
        times.append(time.perf_counter() - s0)
        log(f"batch step {i}: {times[-1] * 1e3:.1f} ms (fields {st.ms_df:.1f} ms, "
            f"pairs {st.ms_hist:.1f} ms)")
    pairs = sum(m.shape[0] * (m.shape[0] - 1) // 2 for m, _ in maps)
    # the same volumes one call each
    s0 = time.perf_counter()
    same = True
    for v in range(nv):
        smap, ids, pst, _ = ctx.ism_volume(vols[v], dims_list[v], isos_list[v], isoamd.EXTRACT_MC,
                                           2, 128, isoamd.DF_LBVH, isoamd.COMPAT_ALL)
        mm = pst.n_nonempty
        same &= bool(np.array_equal(smap.numpy()[: mm * mm].reshape(mm, mm), maps[v][0]))
    loop_s = time.perf_counter() - s0
    ctx.close()
    t = float(np.median(times))
    return {"volumes": nv, "axes": "48..256", "k": "32..128", "ds": 2, "bins": 128,
            "pairs": pairs, "voxels": int(sum(int(np.prod(d)) for d in dims_list)),
            "seconds": round(t, 3), "pairs_per_s": round(pairs / t, 1),
            "pair_launch_ms": round(st.ms_hist, 2), "fields_ms": round(st.ms_df, 2),
            "loop_seconds": round(loop_s, 3), "loop_pairs_per_s": round(pairs / loop_s, 1),
            "same_maps_as_loop": same, "pair_stage": "hip graph",
            "graph_replays": int(st.graph_replays)}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on exactly one GPU; --gpus must be 1 "
                                   "(sharding over several GPUs is not implemented)"}))
        sys.exit(2)
    import isoamd

    isoamd.lib()  # fails loudly when the HIP library is not built
    X, Y, Z = (int(v) for v in a.dims.split(","))
    dims = (X, Y, Z)
    k, ds, H = a.k, a.ds, a.bins
    F = (X // ds) * (Y // ds) * (Z // ds)
    flags = isoamd.COMPAT_ALL if a.mode == "compat" else isoamd.COMPAT_HIST_CHUNKS
    t0 = time.perf_counter()
    img = isoamd.gen_front(X, Y, Z, seed=2)
    isos = np.array([(i + 1) / (k + 1) for i in range(k)], np.float32)
    log(f"volume {dims} generated in {time.perf_counter() - t0:.1f}s")

    ctx = isoamd.Context(0)
    d_vol = ctx.to_device(img)
    smap = ctx.empty((k * k,), np.float32)
    fields = ctx.empty((k, F), np.float32)
    run = lambda method=isoamd.DF_LBVH, iv=isos: ctx.ism_volume(
        d_vol, dims, iv, isoamd.EXTRACT_MC, ds, H, method, flags, fields=fields, smap=smap)

    for i in range(a.warmup):
        _, _, st, _ = run()
        log(f"warmup {i}: m={st.n_nonempty} points={st.total_points} "
            f"extract={st.ms_extract:.1f}ms df={st.ms_df:.1f}ms hist={st.ms_hist:.1f}ms")
    if a.copy_bytes > 0:
        n16 = (a.copy_bytes + 15) // 16
        src, dst = ctx.empty((n16, 4), np.float32), ctx.empty((n16, 4), np.float32)
        ctx.copy(dst, src)
        src.free()
        dst.free()
    ctx.synchronize()
    lat, stages = [], []
    t0 = time.perf_counter()
    for i in range(a.steps):
        s0 = time.perf_counter()
        _, ids, st, _ = run()
        lat.append((time.perf_counter() - s0) * 1e3)
        stages.append((st.ms_extract, st.ms_df, st.ms_hist, st.ms_walk))
        log(f"step {i}: {lat[-1]:.1f} ms")
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    m = st.n_nonempty
    pairs = m * (m - 1) // 2
    value = a.steps * pairs / elapsed
    batched_map = smap.numpy()[: m * m].reshape(m, m).copy()
    # the LBVH fields of the timed map (the brute-force check below reuses the
    # device buffer)
    fields_lbvh = fields.numpy()[:m].copy()
    st_ms = np.mean(np.array(stages), axis=0)

    # plain per-isovalue form
    vbuf = ctx.empty((max(st.total_points // max(m, 1) * 4, 1 << 20), 3), np.float32)
    pfields = ctx.empty((k, F), np.float32)
    plain_ms, plain_same = [], None
    if a.plain_steps:  # untimed: the autotuner measures the per-isovalue shapes here
        plain_ism(ctx, isoamd, d_vol, dims, isos, ds, H, flags, vbuf, pfields)
        ctx.synchronize()
    for i in range(a.plain_steps):
        s0 = time.perf_counter()
        pm, pmap = plain_ism(ctx, isoamd, d_vol, dims, isos, ds, H, flags, vbuf, pfields)
        plain_ms.append((time.perf_counter() - s0) * 1e3)
        log(f"plain {i}: {plain_ms[-1]:.1f} ms")
        plain_same = bool(pm == m and np.array_equal(pmap.numpy(), batched_map))

    # LBVH == brute force on a few isovalues of the same workload
    brute = None
    if a.brute_isos:
        sel = np.linspace(0, k - 1, a.brute_isos).astype(int)
        iv = isos[sel]
        s0 = time.perf_counter()
        _, bids, bst, _ = run(isoamd.DF_BRUTE, iv)
        tb = time.perf_counter() - s0
        fb = fields.numpy()[: bst.n_nonempty]
        want = fields_lbvh[[list(ids).index(s) for s in sel[bids]]]
        # brute force is FP32-VALU bound: one (sample, point) evaluation is
        # 8 FLOP and 7 VALU instructions (3 sub, mul, 2 fma, min; packed FP32
        # is off), so its instruction roof is the lane-operation peak / 7
        evals = float(bst.total_points) * F
        ev_s = evals / (bst.ms_df / 1e3) if bst.ms_df > 0 else 0.0
        roof_ev = VALU_LANES_PER_S / 7.0
        brute = {"isovalues": int(bst.n_nonempty), "lbvh_equals_brute": bool(np.array_equal(fb, want)),
                 "brute_df_ms": bst.ms_df, "wall_s": tb, "points": int(bst.total_points),
                 "roofline": {"bound": "valu", "achieved": round(ev_s * 8 / 1e12, 2),
                              "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": round(ev_s * 8 / 1e12 / FP32_PEAK_TFLOPS, 4),
                              "evaluations_per_s": ev_s,
                              "frac_of_instruction_roof": round(ev_s / roof_ev, 4)}}
        log(f"brute check: {brute}")

    # the LBVH fields: samples/s of the walk, speed-up over brute force per
    # field, and the floor bytes (12 B per point read once + 4 B per sample
    # written) against the HBM roof
    walk_s = st_ms[3] / 1e3
    floor = 12.0 * st.total_points + 4.0 * m * F
    roofline_df = {"kernel": "traverse_kernel (LBVH walk)", "bound": "latency/L2",
                   "walk_ms": round(st_ms[3], 3), "fields_ms": round(st_ms[1], 3),
                   "samples_per_s": (m * F / walk_s) if walk_s > 0 else None,
                   "floor_bytes": floor,
                   "floor_GBs": round(floor / walk_s / 1e9, 1) if walk_s > 0 else None,
                   "floor_frac_of_hbm": round(floor / walk_s / 1e9 / HBM_PEAK_GBS, 4)
                   if walk_s > 0 else None,
                   "speedup_vs_brute_per_field": None}
    if brute and brute["isovalues"] and st_ms[1] > 0:
        roofline_df["speedup_vs_brute_per_field"] = round(
            (brute["brute_df_ms"] / brute["isovalues"]) / (st_ms[1] / m), 1)

    # roofline of the histogram+NMI kernel (algorithmic: 8F bytes per unordered pair)
    default_workload = (dims == (600, 248, 248) and k == 100 and ds == 4 and H == 256
                        and a.mode == "compat")
    hist_s = st_ms[2] / 1e3
    achieved = pairs * 8.0 * F / hist_s / 1e9 if hist_s > 0 else 0.0
    roofline = {"kernel": "hist_nmi_kernel<true> (256 bins)", "bound": "hbm",
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": HIST_TRAFFIC_DEFAULT if default_workload else None,
                "algorithmic_bytes_per_step": pairs * 8 * F}
    if default_workload and hist_s > 0:
        # the counter bytes against the same roof: the last-level cache serves
        # the rest of the algorithmic reads
        roofline["traffic_frac"] = round(HIST_TRAFFIC_DEFAULT / hist_s / 1e9 / HBM_PEAK_GBS, 4)
    if default_workload and walk_s > 0:
        roofline_df["traffic"] = WALK_TRAFFIC_DEFAULT
        roofline_df["traffic_frac"] = round(WALK_TRAFFIC_DEFAULT / walk_s / 1e9 / HBM_PEAK_GBS, 4)

    nmi_ref = None
    cpu = None
    cores = min(16, os.cpu_count() or 1)  # the host-core share of one GPU
    if not a.no_cpu:
        try:
            nmi_ref = nmi_vs_reference(ctx, isoamd, fields, fields_lbvh, batched_map, H, cores)
            log(f"nmi vs reference order: {nmi_ref}")
        except Exception as e:  # the check must not hide the GPU result
            nmi_ref = {"error": repr(e)}
        try:
            cpu = cpu_baseline(img, dims, isos, ds, H, a.mode, cores,
                               gpu_fields=fields_lbvh, gpu_map=batched_map, ids=ids)
        except Exception as e:  # the baseline must not hide the GPU result
            cpu = {"error": repr(e)}

    autotune = ctx.tuning()
    ctx.close()
    batch = None
    if not a.no_batch:
        try:
            batch = batch_bench(a, isoamd)
            log(f"batch: {batch}")
        except Exception as e:  # the batch section must not hide the headline result
            batch = {"error": repr(e)}
    serving = None
    if not a.no_serve:
        try:
            serving = serve_bench(a, isoamd)
        except Exception as e:  # the serving section must not hide the headline result
            serving = {"error": repr(e)}
    p50, p99 = np.percentile(lat, 50), np.percentile(lat, 99)
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "isovalue-pairs/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(elapsed / a.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"{X}x{Y}x{Z} float32 tanh front + 3-octave value noise, "
                                f"{k} isovalues, DF grid {X // ds}x{Y // ds}x{Z // ds} (ds={ds}), "
                                f"LBVH distance fields, {H}-bin joint histograms + NMI"),
                   "form": "batched (iso_ism_volume)", "extractor": "marching cubes",
                   "mode": a.mode, "isosurfaces": m, "points": int(st.total_points),
                   "pairs": pairs},
        "latency_ms": {"p50": round(float(p50), 2), "p99": round(float(p99), 2)},
        "stages_ms": {"extract": round(st_ms[0], 2), "df": round(st_ms[1], 2),
                      "df_walk": round(st_ms[3], 2), "hist_nmi": round(st_ms[2], 2)},
        "plain": ({"value": round(pairs / (np.median(plain_ms) / 1e3), 1),
                   "ms_per_ism": round(float(np.median(plain_ms)), 2), "same_map": plain_same}
                  if plain_ms else None),
        "speedup_vs_plain": (round(float(np.median(plain_ms)) / (elapsed / a.steps * 1e3), 2)
                             if plain_ms else None),
        "brute_check": brute,
        "autotune": autotune,
        "roofline": roofline,
        "roofline_df": roofline_df,
        "nmi_vs_ref32": nmi_ref,
        "cpu_baseline": cpu,
        "batch": batch,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
