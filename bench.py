#!/usr/bin/env python3
"""Benchmark: GTEPS of SSSP and PageRank on RMAT-22 on one MI355X, fraction of
the HBM roofline of the dominant kernel, and batched vs plain SSSP.

`value` is SSSP GTEPS on RMAT scale 22, edge factor 16 (a/b/c/d =
.57/.19/.19/.05, weights U[1,255], seed 3): reached out-arcs / wall time of one
full single-source solve with the graph already resident in HBM.  A "step" is
one such solve.  The PageRank (same graph, pull, d = 0.85, FP64 L1 < 1e-6 or
100 iterations), batched-SSSP, SpMV (RMAT-20), colouring (RMAT-23, symmetrised)
and recursive-BFS (4M-node Zipf tree) figures ride along in the same JSON line,
each with its HBM roofline, its parity check and a CPU time of the oracle, plus
a measured device-to-device copy rate.

The library call is synchronous (it returns after the device finished and the
round counter was read back), so the wall clock around K calls is the device
time plus host launch overhead.
"""
import argparse
import glob
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "show-me-dynamic-parallelism_amd"), os.path.join(ROOT, "oracle")]

HBM_PEAK_GBS = 8000.0
# Autotuner results for this library on MI355X, keyed by arch and degree
# histogram; a miss runs the search and adds the entry.
TUNE_CACHE = os.path.join(ROOT, "show-me-dynamic-parallelism_amd", "tune_cache.json")
METRIC = "GTEPS (SSSP, PageRank) on RMAT-22, 1 MI355X; % HBM roofline; batched vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_for(kernel):
    """HBM bytes per launch of `kernel` from committed rocprofv3 --pmc summaries."""
    for p in glob.glob(os.path.join(ROOT, "profiles", "traffic_*.json")):
        try:
            d = json.load(open(p))
        except (OSError, ValueError):
            continue
        if d.get("kernel") == kernel and "bytes_per_launch" in d:
            return d["bytes_per_launch"]
    return None


def sssp_section(ctx, g, dg, args, smdp, orc, np):
    src = g.source
    sol = smdp.SOL_TUNED
    t0 = time.time()
    params = dg.autotune(smdp.APP_SSSP, src, TUNE_CACHE) if args.tune else ctx.get_params(smdp.APP_SSSP)
    log(f"sssp: params {params.as_dict()} (tuning {time.time() - t0:.1f}s)")
    for _ in range(args.warmup):
        dg.sssp(src, sol)
    t0 = time.perf_counter()
    sts = [dg.sssp(src, sol) for _ in range(args.steps)]
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    st = sts[-1]
    dist = dg.read(smdp.RES_DIST)
    # parity on the full-size graph: exact against Dijkstra
    ok = None
    if not args.skip_check:
        ref = orc.dijkstra(g.row, g.col, g.w, src)
        ok = bool(np.array_equal(dist, ref))
    gteps = st.m_reached / (ms_step * 1e6)
    # roofline pass: per-launch HIP events on the library stream
    ctx.set_profiling(True)
    pst = dg.sssp(src, sol)
    ctx.set_profiling(False)
    bucketed = params.layout == smdp.LAYOUT_BUCKETS or (params.layout == smdp.LAYOUT_AUTO and g.m >= 1 << 20)
    if bucketed:
        # one phase = k_delta_step (tiles) + k_delta_drain (grid chunks); a
        # relaxed arc moves 12 B (col, w, dist[v]), an arc read but left to the
        # bucket's other pass 8 B (col, w), an expanded vertex 20 B (two row
        # offsets, dist[u], queue in/out).  Both kernels form the pass.
        name = "k_delta_step + k_delta_drain"
        bytes_all = 12 * pst.edges_relaxed + 8 * pst.arcs_skipped + 20 * pst.vertices_expanded
        ms_all = pst.ms_dominant + pst.ms_phase2
        dom = {"kernel": name, "bytes": bytes_all, "ms": ms_all, "launches": pst.dominant_launches}
        k1 = {"kernel": "k_delta_step", "bytes": None, "ms": pst.ms_dominant, "launches": pst.dominant_launches}
        k2 = {"kernel": "k_delta_drain", "bytes": None, "ms": pst.ms_phase2, "launches": pst.phase2_launches}
        traffic = _sum_traffic("k_delta_step", "k_delta_drain")
        bytes1, bytes2 = bytes_all, 0
    else:
        e1 = pst.edges_relaxed - pst.edges_phase2
        bytes1 = 12 * e1 + 20 * pst.vertices_expanded  # col + w + dist[v]; 2 offsets + dist[u] + queue in/out
        bytes2 = 12 * pst.edges_phase2
        k1 = {"kernel": "k_expand_tile<false>", "bytes": bytes1, "ms": pst.ms_dominant,
              "launches": pst.dominant_launches}
        k2 = {"kernel": "k_drain<false>", "bytes": bytes2, "ms": pst.ms_phase2, "launches": pst.phase2_launches}
        dom = k1 if k1["ms"] >= k2["ms"] else k2
        traffic = traffic_for(dom["kernel"])
    ach = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9 if dom["ms"] > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": dom["kernel"], "launches": dom["launches"],
            "avg_launch_us": round(dom["ms"] * 1e3 / max(1, dom["launches"]), 3),
            "algorithmic_bytes_per_launch": int(dom["bytes"] / max(1, dom["launches"]))}
    kernels = []
    for k in (k1, k2):
        entry = {"kernel": k["kernel"], "ms_total": round(k["ms"], 4), "launches": k["launches"]}
        if k["bytes"] is not None:
            a = k["bytes"] / (k["ms"] * 1e-3) / 1e9 if k["ms"] > 0 else 0.0
            entry.update(achieved_GBs=round(a, 1), frac=round(a / HBM_PEAK_GBS, 4))
        kernels.append(entry)
    whole = (bytes1 + bytes2) / (ms_step * 1e-3) / 1e9
    inflation = st.edges_relaxed / max(1, st.m_reached)
    roof["frac_min_bytes"] = round(roof["frac"] / max(inflation, 1e-9), 4)
    # end to end: the drop-in call with host arrays (upload, solve, download)
    e2e = ctx.sssp(g, src, sol)[1]
    dev = sorted(x.ms_kernel for x in sts)
    info = {"source": int(src), "ms_per_solve": round(ms_step, 4), "ms_device": round(st.ms_kernel, 4),
            "ms_device_median": round(dev[len(dev) // 2], 4),
            "ms_end_to_end_dropin": round(e2e.ms_total, 3),
            "gteps_end_to_end_dropin": round(st.m_reached / (e2e.ms_total * 1e6), 4),
            "rounds": st.rounds, "edges_relaxed": st.edges_relaxed, "m_reached": st.m_reached,
            "inflation": round(st.edges_relaxed / max(1, st.m_reached), 3),
            "vertices_expanded": st.vertices_expanded, "kernel_launches": st.kernel_launches,
            "exact_vs_dijkstra": ok, "atomics_tried": st.atomics_tried, "atomics_won": st.atomics_won,
            "arcs_skipped": st.arcs_skipped, "whole_solve_GBs": round(whole, 1), "kernels": kernels,
            "layout": "buckets" if bucketed else "unordered rounds", "params": params.as_dict()}
    return gteps, ms_step, roof, info, st


def cpu_baseline(g, src, orc, np, budget_s=8.0):
    """The oracle's worklist Bellman-Ford (the reference algorithm) on all host
    threads, with the single-thread run and a binary-heap Dijkstra beside it."""
    def timed(fn, budget, cap):
        runs, t0, out = 0, time.perf_counter(), None
        while True:
            out = fn()
            runs += 1
            if time.perf_counter() - t0 >= budget or runs >= cap:
                break
        return (time.perf_counter() - t0) / runs, runs, out
    mr = None
    t_par, n_par, (d, rounds, edges, threads) = timed(lambda: orc.sssp_worklist_par(g.row, g.col, g.w, src), budget_s, 20)
    mr = int(np.diff(g.row)[d != orc.INF].sum())
    t_seq, n_seq, (_, _, _) = timed(lambda: orc.sssp_worklist(g.row, g.col, g.w, src), budget_s, 20)
    t_dij, n_dij, _ = timed(lambda: orc.dijkstra(g.row, g.col, g.w, src), budget_s / 2, 5)
    return {"value": round(mr / t_par / 1e9, 5), "unit": "GTEPS", "cores": threads, "kind": "port",
            "sample": f"{n_par} full solves of worklist Bellman-Ford on {threads} threads (sssp_kernel.cu:40-61 "
                      f"restated, oracle/oracle.c orc_sssp_worklist_par) on the same RMAT-22 graph and source; "
                      f"{t_par:.3f} s each, {rounds} rounds, {edges} arcs scanned",
            "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "single_thread": {"value": round(mr / t_seq / 1e9, 5), "cores": 1, "s_per_solve": round(t_seq, 3),
                              "solves": n_seq},
            "dijkstra_binary_heap": {"value": round(mr / t_dij / 1e9, 5), "cores": 1, "s_per_solve": round(t_dij, 3),
                                     "solves": n_dij}}


def c1_section(ctx, smdp, orc, np):
    """The reference's own small case: 16384 vertices, 8 uniform out-arcs each,
    weights U[1,100], source 0, written as a DIMACS9 .gr file and read back
    with format 0 (sssp.cpp:100-154, global.cpp:29-64), solved on the GPU
    through the drop-in call and through the gpu-sssp command, with the
    oracle's worklist Bellman-Ford timed on the host beside it."""
    import subprocess
    import tempfile
    src_g = smdp.gen_uniform(16384, 8, 1, 100, 1)
    tmp = tempfile.mkdtemp()
    path = os.path.join(tmp, "uniform16k.gr")
    smdp.write_dimacs9(path, src_g)
    t_read = []
    for _ in range(5):
        t0 = time.perf_counter()
        g = smdp.read_graph(path, 0)
        t_read.append(time.perf_counter() - t0)
    same = bool(np.array_equal(g.row, src_g.row) and np.array_equal(g.col, src_g.col)
                and np.array_equal(g.w, src_g.w) and g.source == src_g.source)
    ref_read_ms = None
    if orc.ref_available():
        t0 = time.perf_counter()
        rg = orc.ref_read(path, 0)
        ref_read_ms = round((time.perf_counter() - t0) * 1e3, 2)
        same = same and bool(np.array_equal(rg["row"], g.row) and np.array_equal(rg["col"], g.col)
                             and np.array_equal(rg["w"], g.w))
    src = g.source
    sol = smdp.SOL_TUNED
    for _ in range(3):
        ctx.sssp(g, src, sol)
    runs = [ctx.sssp(g, src, sol) for _ in range(21)]
    e2e = sorted(st.ms_total for _, st in runs)[10]
    dev = sorted(st.ms_kernel for _, st in runs)[10]
    dist, st = runs[-1]
    ref = orc.dijkstra(g.row, g.col, g.w, src)
    # CPU: the reference algorithm restated, one thread, about 2 s of solves
    n_cpu, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 2.0:
        d_cpu, rounds_cpu, _ = orc.sssp_worklist(g.row, g.col, g.w, src)
        n_cpu += 1
    t_cpu = (time.perf_counter() - t0) / n_cpu
    cli = {}
    exe = os.path.join(ROOT, "show-me-dynamic-parallelism_amd", "bin", "gpu-sssp")
    if os.path.exists(exe):
        r = subprocess.run([exe, "-i", path, "-f", "0", "-s", "6", "-v"], capture_output=True, text=True,
                           timeout=120)
        for key, pat in (("s_kernel", "HIP kernel execution"), ("s_total", "HIP total"),
                         ("s_read", "Read data and AdjList to CSR")):
            for ln in (r.stdout + r.stderr).splitlines():
                if ln.startswith(pat):
                    cli[key] = float(ln.split()[-1])
        cli["returncode"] = r.returncode
    os.remove(path)
    os.rmdir(tmp)
    return {"workload": "uniform16k .gr, 16384 vertices x 8 arcs, weights U[1,100], source 0",
            "file_read_ms": round(sorted(t_read)[2] * 1e3, 2), "reference_reader_ms": ref_read_ms,
            "read_back_identical": same, "gpu_ms_device": round(dev, 4), "gpu_ms_end_to_end_dropin": round(e2e, 3),
            "rounds": st.rounds, "exact_vs_dijkstra": bool(np.array_equal(dist, ref)),
            "cpu": {"ms": round(t_cpu * 1e3, 3), "cores": 1, "kind": "port", "rounds": rounds_cpu,
                    "exact": bool(np.array_equal(d_cpu, ref)),
                    "sample": f"{n_cpu} solves of worklist Bellman-Ford (oracle orc_sssp_worklist)"},
            "gpu_vs_cpu_end_to_end": round(t_cpu * 1e3 / max(e2e, 1e-9), 2), "cli_gpu_sssp_f0": cli}


def pagerank_section(ctx, g, args, smdp, orc, np):
    ig = smdp.invert(g)
    od = g.outdegree()
    opts = smdp.pr_opts(0.85, 1.0, 1e-6, 100)
    with smdp.Graph(ctx, smdp.HostGraph(ig.row, ig.col)) as dg:
        dg.set_outdeg(od)
        sol = smdp.SOL_TUNED
        if args.tune:
            dg.autotune(smdp.APP_PAGERANK, 0, TUNE_CACHE)
        for _ in range(max(1, args.warmup)):
            dg.pagerank(opts, sol)
        t0 = time.perf_counter()
        sts = [dg.pagerank(opts, sol) for _ in range(args.steps)]
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        st = sts[-1]
        r = dg.read(smdp.RES_RANK, np.float32)
        ctx.set_profiling(True)
        pst = dg.pagerank(opts, sol)
        ctx.set_profiling(False)
    iters = int(st.rounds)
    gteps = g.m * iters / (ms * 1e6)
    per_iter = 8 * g.m + 20 * g.n
    params = ctx.get_params(smdp.APP_PAGERANK)
    binned = params.layout == smdp.LAYOUT_BINNED or (params.layout == smdp.LAYOUT_AUTO and g.m >= 1 << 20)
    # every launch after the stop rule fired returns at once: bytes of the
    # iterations run, over the summed duration of all launches.  The binned
    # iteration is two kernels (scatter, reduce); the roofline covers both.
    if binned:
        kernel = "k_bin_scatter + k_bin_reduce"
        ms_k = pst.ms_dominant + pst.ms_phase2
        launches = pst.dominant_launches
        tr = [traffic_for("k_bin_scatter"), traffic_for("k_bin_reduce")]
        traffic = None if None in tr else tr[0] + tr[1]
    else:
        kernel = "k_rows_tuned<true>"
        ms_k = pst.ms_dominant
        launches = pst.dominant_launches
        traffic = traffic_for(kernel)
    ach = per_iter * iters / (ms_k * 1e-3) / 1e9 if ms_k > 0 else 0.0
    out = {"gteps": round(gteps, 3), "ms_per_solve": round(ms, 4), "iterations": iters,
           "residual": st.residual, "layout": "binned" if binned else params.layout,
           "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "kernel": kernel,
                        "traffic": traffic, "launches": launches,
                        "avg_launch_us": round(ms_k * 1e3 / max(1, launches), 3),
                        "algorithmic_bytes_per_launch": int(per_iter * iters / max(1, launches))},
           "kernels": [{"kernel": "k_bin_reduce" if binned else kernel, "ms_total": round(pst.ms_dominant, 4),
                        "launches": pst.dominant_launches},
                       {"kernel": "k_bin_scatter" if binned else "phase 2", "ms_total": round(pst.ms_phase2, 4),
                        "launches": pst.phase2_launches}]}
    if not args.skip_check:
        t0 = time.perf_counter()
        ref, _, _ = orc.pagerank(ig.row, ig.col, od, d=0.85, T=1.0, fixed_iters=iters)
        t_seq = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, _, nt = orc.pagerank_par(ig.row, ig.col, od, iters, d=0.85, T=1.0)
        t_par = time.perf_counter() - t0
        out["cpu"] = _cpu_entry(t_seq, t_par, nt, g.m * iters,
                                f"{iters} FP64 pull iterations (pagerank_kernel.cu:95-120 restated, "
                                "orc_pagerank_f64 / orc_pagerank_f64_par) on the same in-CSR")
        out["l1_vs_fp64_iteration_matched"] = float(np.abs(r.astype(np.float64) - ref).sum())
        out["linf_vs_fp64"] = float(np.abs(r.astype(np.float64) - ref).max())
        conv, it_conv, _ = orc.pagerank(ig.row, ig.col, od, d=0.85, T=1.0, tol=1e-9, max_iter=1000)
        out["l1_vs_fp64_converged"] = float(np.abs(r.astype(np.float64) - conv).sum())
        out["fp64_converged_iterations"] = it_conv
    return out


def batched_section(ctx, args, smdp, orc, np):
    """64 sources on RMAT-21 in one captured hipGraph (one persistent launch
    runs all rounds) against 64 plain solves on the same resident graph, and
    256 small graphs in one captured launch against 256 plain solves."""
    out = {}
    g = smdp.gen_rmat(args.batch_scale, 16, wlo=1, whi=255, seed=4)
    srcs = smdp.gen_sources(g, 64, 4)
    with smdp.Graph(ctx, g) as dg:
        dg.sssp_multi_source(srcs)  # capture and warm-up
        sts = [dg.sssp_multi_source(srcs) for _ in range(10)]
        ms_batch = sorted(x.ms_kernel for x in sts)[5]
        st = sts[-1]
        ctx.set_profiling(True)
        pst = dg.sssp_multi_source(srcs)
        ctx.set_profiling(False)
        for s in srcs[:2]:
            dg.sssp(int(s), smdp.SOL_TUNED)
        plain = []
        m_reached = 0
        for s in srcs:
            x = dg.sssp(int(s), smdp.SOL_TUNED)
            plain.append(x.ms_kernel)
            m_reached += x.m_reached
        t_plain = sum(plain)
    # an arc visit reads col and w (8 B) and the 64 distances of its target
    # (one 256-B row): 264 B; the CSR bytes are shared by the 64 sources
    per_visit = 8 + 4 * 64
    ms_k = pst.ms_dominant
    roof = _roof(per_visit * pst.edges_relaxed, ms_k)
    roof.update(kernel="k_ms_run (hipGraph replay)", launches=pst.dominant_launches,
                avg_launch_us=round(ms_k * 1e3 / max(1, pst.dominant_launches), 3), traffic=traffic_for("k_ms_run"))
    entry = {"graph": f"rmat-{args.batch_scale} ef16 w[1,255] seed 4", "sources": 64,
             "ms_batched_median_of_10": round(ms_batch, 3), "ms_64_plain": round(t_plain, 3),
             "speedup": round(t_plain / ms_batch, 2), "rounds": st.rounds, "arc_visits": st.edges_relaxed,
             "gteps": round(m_reached / (ms_batch * 1e6), 3), "roofline": roof}
    if not args.skip_cpu:
        # the oracle's Dijkstra (the strongest sequential baseline), 16 of the
        # 64 sources on 16 host threads, scaled to the 64
        from concurrent.futures import ThreadPoolExecutor
        sample = [int(s) for s in srcs[:16]]
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=16) as ex:
            list(ex.map(lambda s: orc.dijkstra(g.row, g.col, g.w, s), sample))
        t16 = time.perf_counter() - t0
        t64 = t16 * 64 / len(sample)
        entry["cpu"] = {"ms": round(t64 * 1e3, 1), "cores": 16, "kind": "port",
                        "gteps": round(m_reached / t64 / 1e9, 4),
                        "sample": "binary-heap Dijkstra (orc_dijkstra) from 16 of the 64 sources, one per host "
                                  "thread, scaled to 64"}
    out["multi_source"] = entry
    graphs = [smdp.gen_small_graph(i, seed=4) for i in range(256)]
    srcs = np.zeros(256, dtype=np.int32)
    with smdp.Batch(ctx, graphs, srcs) as b:
        plan = b.autotune() if args.tune else ctx.get_params(smdp.APP_MSSSP)
        b.run()
        sts = [b.run() for _ in range(10)]
        t_batch = sorted(x.ms_kernel for x in sts)[5]
        st = sts[-1]
    t0 = time.perf_counter()
    for gi in graphs:
        ctx.sssp(gi, 0, smdp.SOL_TUNED)
    t_plain_upload = time.perf_counter() - t0
    devs = [smdp.Graph(ctx, gi) for gi in graphs]
    for d in devs[:4]:
        d.sssp(0, smdp.SOL_TUNED)
    t_plain = sum(d.sssp(0, smdp.SOL_TUNED).ms_kernel for d in devs)
    for d in devs:
        d.close()
    out["multi_graph"] = {"graphs": 256, "vertices": int(sum(x.n for x in graphs)),
                          "arcs": int(sum(x.m for x in graphs)), "ms_batched_median_of_10": round(t_batch, 3),
                          "edges_relaxed": st.edges_relaxed,
                          "roofline": dict(_roof(12 * st.edges_relaxed + 20 * st.vertices_expanded, t_batch),
                                           kernel="k_mg (hipGraph replay)", traffic=traffic_for("k_mg")),
                          "plan": {"wg_per_cu": plan.wg_per_cu, "group_share": plan.group_share,
                                   "autotuned": bool(args.tune)},
                          "ms_256_plain_resident": round(t_plain, 3),
                          "ms_256_plain_with_upload": round(t_plain_upload * 1e3, 3),
                          "speedup": round(t_plain / t_batch, 2)}
    return out


def _median_runs(fn, reps=5):
    fn()
    sts = [fn() for _ in range(reps)]
    ms = sorted(x.ms_kernel for x in sts)
    return ms[len(ms) // 2], sts[-1]


def _cpu_entry(t_seq, t_par, threads, arcs, sample):
    """CPU baseline of one section: the oracle on one thread and on all host threads."""
    return {"ms": round(t_seq * 1e3, 1), "cores": 1, "kind": "port",
            "ms_parallel": round(t_par * 1e3, 1), "threads": threads,
            "gteps": round(arcs / t_seq / 1e9, 4), "gteps_parallel": round(arcs / t_par / 1e9, 4),
            "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "sample": sample}


def _sum_traffic(*kernels):
    t = [traffic_for(k) for k in kernels]
    return None if any(x is None for x in t) else int(sum(t))


def _roof(bytes_, ms):
    ach = bytes_ / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "algorithmic_bytes": int(bytes_)}


def spmv_section(ctx, smdp, orc, np):
    """CSR SpMV y = A x, RMAT scale 20 (16.8M nnz), FP32 with compensated sums."""
    g = smdp.gen_rmat(20, 16, seed=2)
    val = smdp.gen_uniform_f32(g.m, -1.0, 1.0, 2, 6)
    x = smdp.gen_uniform_f32(g.n, 0.0, 1.0, 2, 9)
    with smdp.Graph(ctx, g) as dg:
        dg.set_values(val)
        dg.set_x(x)
        ms, st = _median_runs(lambda: dg.spmv(smdp.SOL_TUNED))
        y = dg.read(smdp.RES_Y, np.float32)
    t0 = time.perf_counter()
    y64 = orc.spmv(g.row, g.col, val, x)
    t_cpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, nt = orc.spmv_par(g.row, g.col, val, x)
    t_par = time.perf_counter() - t0
    return {"workload": "spmv_rmat20", "rows": g.n, "nnz": g.m, "ms_device_median": round(ms, 4),
            "gflops": round(2 * g.m / (ms * 1e6), 2), "gteps": round(g.m / (ms * 1e6), 3),
            "roofline": dict(_roof(12 * g.m + 8 * g.n + 4, ms), kernel="k_permute_x + k_spmv_merge + k_merge_fixup",
                             traffic=_sum_traffic("k_permute_x", "k_spmv_merge", "k_merge_fixup")),
            "linf_vs_fp64": float(np.abs(y.astype(np.float64) - y64).max()),
            "cpu": _cpu_entry(t_cpu, t_par, nt, g.m, "one thread-per-row FP64 product (SpMV_kernel.cu:81-96 "
                              "restated, orc_spmv_f64 / orc_spmv_f64_par)")}


def color_section(ctx, smdp, orc, np):
    """Colouring of RMAT scale 23, edge factor 8, symmetrised, random priorities."""
    g = smdp.gen_rmat(23, 8, symmetrize=True, seed=5)
    prio = smdp.gen_prio(g.n, 5)
    with smdp.Graph(ctx, g) as dg:
        dg.set_prio(prio)
        ms, st = _median_runs(lambda: dg.color(True, smdp.SOL_TUNED), reps=3)
        c = dg.read(smdp.RES_COLOR)
    t0 = time.perf_counter()
    ref, k = orc.color_closed(g.row, g.col, prio)
    t_cpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, _, nt = orc.color_layers_par(g.row, g.col, prio)
    t_par = time.perf_counter() - t0
    return {"workload": "color_rmat23_sym", "vertices": g.n, "arcs": g.m, "ms_device_median": round(ms, 4),
            "rounds": int(st.rounds), "colors": int(c.max()), "exact_vs_oracle": bool(np.array_equal(c, ref)),
            "scanned_arcs": int(st.edges_relaxed), "gteps": round(g.m / (ms * 1e6), 3),
            "roofline": dict(_roof(12 * st.edges_relaxed + 16 * st.vertices_expanded, ms),
                             traffic_per_launch={"k_layer_short": traffic_for("k_layer_short")}),
            "cpu": _cpu_entry(t_cpu, t_par, nt, g.m,
                              "the closed form colour(v) = 1 + max over preceding neighbours: one pass in "
                              "precedence order (orc_color_closed) / by layers (orc_color_layers_par); the round "
                              "rule of graph_color_kernel.cu:28-49 gives the same colours")}


def bfs_section(ctx, smdp, orc, np):
    """Recursive-BFS workload: a 4M-node tree with Zipf(1.5) child counts capped at 4096."""
    t = smdp.gen_zipf_tree(4194304, 1.5, 4096, 5)
    with smdp.Graph(ctx, t) as dg:
        ms, st = _median_runs(lambda: dg.bfs(0, smdp.SOL_TUNED))
        lv = dg.read(smdp.RES_LEVEL)
    t0 = time.perf_counter()
    ref = orc.bfs(t.row, t.col, 0)
    t_cpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, nt = orc.bfs_par(t.row, t.col, 0)
    t_par = time.perf_counter() - t0
    return {"workload": "bfs_zipf_tree_4m", "vertices": t.n, "arcs": t.m, "ms_device_median": round(ms, 4),
            "rounds": int(st.rounds), "exact_vs_oracle": bool(np.array_equal(lv, ref)),
            "gteps": round(t.m / (ms * 1e6), 3),
            "roofline": dict(_roof(8 * st.edges_relaxed + 12 * t.n, ms),
                             traffic_per_launch={"k_drain<true>": traffic_for("k_drain<true>")}),
            "cpu": _cpu_entry(t_cpu, t_par, nt, t.m, "one BFS from the root: FIFO (bfs_rec.cpp:102-121 restated, "
                              "orc_bfs_fifo) / level-synchronous (orc_bfs_par)")}


def tree_section(ctx, smdp, orc, np):
    """Tree Height / Tree Descendant (named in the reference's README.md:12-13)
    on the same 4M-node Zipf(1.5) tree: a down pass per level, then an up pass."""
    t = smdp.gen_zipf_tree(4194304, 1.5, 4096, 5)
    with smdp.Graph(ctx, t) as dg:
        ms, st = _median_runs(lambda: dg.tree(0))
        h = dg.read(smdp.RES_HEIGHT)
        d = dg.read(smdp.RES_DESC)
    t0 = time.perf_counter()
    rh, rd, rth = orc.tree(t.row, t.col, 0)
    t_cpu = time.perf_counter() - t0
    # down: 8 B/arc (child id read, slot write) + 12 B/vertex (slot read, two row offsets);
    # up: 12 B/arc (child id, child height and count) + 20 B/vertex (slot, two offsets, two results)
    alg = 20 * t.m + 32 * t.n
    return {"workload": "tree_zipf_4m", "vertices": t.n, "arcs": t.m, "ms_device_median": round(ms, 4),
            "levels": int(st.rounds), "tree_height": int(rth), "descendants_of_root": int(rd[0]),
            "exact_vs_oracle": bool(np.array_equal(h, rh) and np.array_equal(d, rd)),
            "gteps": round(t.m / (ms * 1e6), 3),
            "roofline": dict(_roof(alg, ms), traffic_per_launch={k: traffic_for(k) for k in ("k_tree_down", "k_tree_up")}),
            "cpu": {"ms": round(t_cpu * 1e3, 1), "cores": 1, "kind": "port", "gteps": round(t.m / t_cpu / 1e9, 4),
                    "sample": "one pass: BFS order from the root, then that order backwards (orc_tree)"}}


def stream_copy_bandwidth(nbytes=1 << 29, reps=5):
    """Measured device-to-device copy rate (read + write bytes per second)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    hip.hipEventSynchronize.argtypes = [C.c_void_p]
    hip.hipEventDestroy.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    a, b, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
    best = 1e30
    for _ in range(reps):
        assert hip.hipEventRecord(e0, None) == 0
        assert hip.hipMemcpy(b, a, nbytes, 3) == 0  # hipMemcpyDeviceToDevice
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = C.c_float()
        assert hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
        best = min(best, ms.value)
    for x in (a, b):
        hip.hipFree(x)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    return round(2 * nbytes / (best * 1e-3) / 1e9, 1)


def copy_calibration(nbytes=1 << 28, reps=4):
    """A plain device-to-device copy of known size (hipMemcpy), so that
    rocprofv3 FETCH_SIZE / WRITE_SIZE readings can be scaled to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {reps} copies of {nbytes} bytes")


def only_section(args, smdp, orc, np):
    ctx = smdp.Context(0)
    ctx.set_params(smdp.APP_PAGERANK, layout=args.layout)
    if args.params:
        kw = json.loads(args.params)
        app = {"sssp": smdp.APP_SSSP, "pagerank": smdp.APP_PAGERANK, "spmv": smdp.APP_SPMV,
               "color": smdp.APP_COLOR, "bfs": smdp.APP_BFS}.get(args.only, smdp.APP_MSSSP)
        ctx.set_params(app, **kw)
    if args.only in ("sssp", "pagerank"):
        g = smdp.gen_rmat(args.scale, 16, wlo=1, whi=255, seed=3)
        if args.only == "sssp":
            with smdp.Graph(ctx, g) as dg:
                if args.tune and not args.params:
                    dg.autotune(smdp.APP_SSSP, g.source, TUNE_CACHE)
                for _ in range(args.warmup + args.steps):
                    st = dg.sssp(g.source, smdp.SOL_TUNED)
                log(f"sssp: {st.ms_kernel:.3f} ms, {st.rounds} rounds, {st.edges_relaxed} arcs, "
                    f"params {ctx.get_params(smdp.APP_SSSP).as_dict()}")
                if args.round_log:
                    ctx.set_profiling(True)
                    dg.sssp(g.source, smdp.SOL_TUNED)
                    q, arcs, ms = ctx.round_log()
                    ctx.set_profiling(False)
                    modes = ("sweep-light", "light", "settle", "done", "settle-pull")
                    for i in range(len(q)):
                        if q[i] or arcs[i]:
                            # bucketed phases log mode << 60 | bucket << 32 | entries
                            tag = (f"{modes[q[i] >> 60]} bucket {(q[i] >> 32) & 0xFFFFFFF} "
                                   if ctx.get_params(smdp.APP_SSSP).layout == smdp.LAYOUT_BUCKETS else "")
                            log(f"  round {i}: {tag}queue {q[i] & 0xFFFFFFFF} arcs {arcs[i]} {ms[i]:.3f} ms "
                                f"{arcs[i] / max(ms[i], 1e-9) / 1e6:.1f} Garcs/s")
        else:
            ig = smdp.invert(g)
            with smdp.Graph(ctx, smdp.HostGraph(ig.row, ig.col)) as dg:
                dg.set_outdeg(g.outdegree())
                if args.tune and not args.params and args.layout == 0:
                    dg.autotune(smdp.APP_PAGERANK, 0, TUNE_CACHE)
                t0 = time.time()
                for _ in range(args.warmup + args.steps):
                    st = dg.pagerank(smdp.pr_opts(0.85, 1.0, 1e-6, 100), smdp.SOL_TUNED)
                ctx.set_profiling(True)
                pst = dg.pagerank(smdp.pr_opts(0.85, 1.0, 1e-6, 100), smdp.SOL_TUNED)
                ctx.set_profiling(False)
                log(f"pagerank: {st.ms_kernel:.3f} ms, {st.rounds} iterations, kernels {pst.ms_dominant:.3f} ms "
                    f"({pst.dominant_launches} launches) + {pst.ms_phase2:.3f} ms ({pst.phase2_launches}), "
                    f"params {ctx.get_params(smdp.APP_PAGERANK).as_dict()}, {time.time() - t0:.1f} s with set-up")
    elif args.only == "spmv":
        g = smdp.gen_rmat(20, 16, seed=2)
        with smdp.Graph(ctx, g) as dg:
            dg.set_values(smdp.gen_uniform_f32(g.m, -1.0, 1.0, 2, 6))
            dg.set_x(smdp.gen_uniform_f32(g.n, 0.0, 1.0, 2, 9))
            for _ in range(args.warmup + args.steps):
                st = dg.spmv(smdp.SOL_TUNED)
            log(f"spmv rmat-20: {st.ms_kernel:.4f} ms, {st.edges_relaxed} nnz")
    elif args.only == "color":
        g = smdp.gen_rmat(23, 8, symmetrize=True, seed=5)
        with smdp.Graph(ctx, g) as dg:
            dg.set_prio(smdp.gen_prio(g.n, 5))
            for _ in range(args.warmup + args.steps):
                st = dg.color(True, smdp.SOL_TUNED)
            log(f"colour rmat-23: {st.ms_kernel:.3f} ms, {st.rounds} rounds")
    elif args.only == "bfs":
        t = smdp.gen_zipf_tree(4194304, 1.5, 4096, 5)
        with smdp.Graph(ctx, t) as dg:
            for _ in range(args.warmup + args.steps):
                st = dg.bfs(0, smdp.SOL_TUNED)
            log(f"bfs zipf tree 4M: {st.ms_kernel:.3f} ms, {st.rounds} rounds")
    elif args.only == "c1":
        log(json.dumps(c1_section(ctx, smdp, orc, np)))
    elif args.only == "tree":
        t = smdp.gen_zipf_tree(4194304, 1.5, 4096, 5)
        with smdp.Graph(ctx, t) as dg:
            for _ in range(args.warmup + args.steps):
                st = dg.tree(0)
            log(f"tree zipf 4M: {st.ms_kernel:.3f} ms, {st.rounds} levels")
    else:
        log(json.dumps(batched_section(ctx, args, smdp, orc, np)))
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--batch-scale", type=int, default=21)
    ap.add_argument("--no-tune", dest="tune", action="store_false")
    ap.add_argument("--skip-check", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-pagerank", action="store_true")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--skip-extra", action="store_true", help="skip the SpMV, colouring and BFS sections")
    ap.add_argument("--layout", type=int, default=0,
                    help="PageRank arc layout: 0 auto, 1 plain CSR, 2 binned, 8 XCD blocks")
    ap.add_argument("--params", default="", help="profiling runs: JSON object of smdp_params fields to set")
    ap.add_argument("--round-log", action="store_true", help="profiling runs: print SSSP per-round log")
    ap.add_argument("--only", choices=["sssp", "pagerank", "batched", "copy", "spmv", "color", "bfs", "tree", "c1"],
                    help="profiling runs: one section only, no JSON contract line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build measures one GPU "
              f"(the multi-GPU sharding is design only).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import orc
    import smdp

    if args.only == "copy":
        copy_calibration()
        return
    if args.only:
        only_section(args, smdp, orc, np)
        return

    t0 = time.time()
    g = smdp.gen_rmat(args.scale, 16, wlo=1, whi=255, seed=3)
    log(f"graph: rmat-{args.scale} n={g.n} m={g.m} source={g.source} ({time.time() - t0:.1f}s)")
    ctx = smdp.Context(0)
    info_dev = ctx.info()
    dg = smdp.Graph(ctx, g)
    gteps, ms_step, roof, sssp_info, st = sssp_section(ctx, g, dg, args, smdp, orc, np)
    log(f"sssp: {gteps:.3f} GTEPS, {ms_step:.3f} ms/solve, roofline {roof['frac']:.3f}")
    dg.close()

    line = {"metric": METRIC, "value": round(gteps, 4), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "int32",
            "data": f"synthetic: seeded RMAT scale {args.scale}, edge factor 16, a/b/c/d=.57/.19/.19/.05, "
                    f"vertex ids permuted, weights U[1,255], seed 3",
            "config": {"workload": f"sssp_rmat{args.scale}", "scale": args.scale, "edge_factor": 16,
                       "vertices": g.n, "arcs": g.m, "solution": "tuned", "device": info_dev},
            "roofline": roof, "sssp": sssp_info}
    if not args.skip_cpu:
        line["cpu_baseline"] = cpu_baseline(g, g.source, orc, np)
        log(f"cpu baseline: {line['cpu_baseline']['value']} GTEPS")
    try:
        line["c1"] = c1_section(ctx, smdp, orc, np)
        log(f"c1: {line['c1']}")
    except smdp.SmdpError as e:
        line["c1"] = {"error": str(e)}
    if not args.skip_pagerank:
        try:
            line["pagerank"] = pagerank_section(ctx, g, args, smdp, orc, np)
            log(f"pagerank: {line['pagerank']}")
        except smdp.SmdpError as e:
            line["pagerank"] = {"error": str(e)}
    if not args.skip_batched:
        try:
            line["batched"] = batched_section(ctx, args, smdp, orc, np)
            log(f"batched: {line['batched']}")
        except smdp.SmdpError as e:
            line["batched"] = {"error": str(e)}
    if not args.skip_extra:
        for name, fn in (("spmv", spmv_section), ("color", color_section), ("bfs", bfs_section),
                         ("tree", tree_section)):
            try:
                line[name] = fn(ctx, smdp, orc, np)
                log(f"{name}: {line[name]}")
            except smdp.SmdpError as e:
                line[name] = {"error": str(e)}
    line["stream_copy_GBs"] = stream_copy_bandwidth()
    log(f"stream copy: {line['stream_copy_GBs']} GB/s")
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
