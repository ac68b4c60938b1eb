"""Benchmark on one MI355X.

Headline (value): plain BFS + SSSP on R-MAT scale 22 (undirected, edge factor
16, int32 weights in [1, 64], source = highest-degree vertex).  A step is one
BFS and one SSSP from the source, graph resident in HBM and outputs kept on
the device; value = GTEPS over the timed steps, counting the sum of CSR
out-degrees of the reached vertices per traversal (the reference's counting,
Algorithms/BFS/bfs.cu:963-980); the undirected-edge count is half of it.
The graph's options are first picked by gg_autotune (BFS + SSSP, outside
the timed region; --no-tune keeps the defaults).  `value` comes from a timed
region without per-kernel events; a second timed pass of the same steps
records HIP events around every kernel on the library's stream
(GG_INSTRUMENT), and `roofline` is that of the dominant kernel class, with
its frontier-heavy launches (>= 1% of m edges) broken out.  `cpu_baseline` is the oracle (single-threaded C restatement of the
reference's CPUBFS and, for SSSP, a heap Dijkstra with CPUSSSP's output) on
the same graph and source.

Per-config rows (same JSON line, `per_config`), each with the GPU time, its
GTEPS (the reference's count, the undirected half where the graph is
symmetric, and work GTEPS = edges inspected / t), the oracle as written on
1 host core and the OpenMP baselines on all host cores, and a parity check
of the GPU result against the oracle in the same run:
  1. BFS on the Erdos-Renyi graph (n = 4096);
  2. the headline's BFS and SSSP, plus CC on the same graph;
  3. CC on the 2048^2 percolated lattice (label propagation, hooking, the
     generic GAS program);
  4. batched 128 SSSP + 64 PPR on R-MAT scale 20 directed against 128 + 64
     plain GPU launches and against the CPU's per-instance loop (a sample of
     instances, scaled);
  5. PageRank and its adjoint on R-MAT scale 22 directed with FP32 weights,
     with the L1 error against the FP64 oracle in every run.
The generic GAS path (the reference's own schedules: push-only BFS, GAS
SSSP, GAS CC) is timed beside the specialised kernels.  The R-MAT graphs are
stored renumbered by degree (GG_GRAPH_REORDER; ids in and out stay the
generator's; --no-reorder keeps the generated order), and the headline's
BFS / SSSP times on the graph in generated order are reported beside it.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# `roofline.traffic`: HBM bytes per launch of a kernel class, from the
# committed rocprofv3 --pmc summary (a table row per class: FETCH_SIZE x2 +
# WRITE_SIZE x1, calibrated by a 1 GiB copy, per launch of the instrumented
# pass).
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "headline_rmat22_hbm_traffic_pmc.md")


def measured_traffic(kernel_class):
    """HBM bytes per launch of a kernel class (fetch + write per launch of the
    instrumented pass) from the committed counter summary, or None."""
    if not os.path.exists(TRAFFIC_SUMMARY):
        return None
    for line in open(TRAFFIC_SUMMARY):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[1] == kernel_class:
            try:
                return {"bytes": (float(cells[4]) + float(cells[5])) * 1e6, "kernel": cells[0],
                        "source": "profiles/" + os.path.basename(TRAFFIC_SUMMARY)}
            except ValueError:
                return None
    return None


def median_ms(fn, reps):
    import numpy as np

    fn()
    ts = []
    st = None
    for _ in range(reps):
        _, st = fn()[:2]
        ts.append(st["gpu_ms"])
    return float(np.median(ts)), st


def headline(args, gg, _abi, workloads, torch, np):
    t0 = time.perf_counter()
    csr = workloads.config2_rmat(args.scale)
    src = workloads.max_degree_vertex(csr)
    log(f"R-MAT s{args.scale}: n={csr.n} m={csr.m} src={src} (deg {csr.degrees()[src]}) "
        f"built in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda:0")
    out_bfs = torch.empty(csr.n, dtype=torch.int32, device=dev)
    out_sssp = torch.empty(csr.n, dtype=torch.int32, device=dev)
    t0 = time.perf_counter()
    g = gg.Graph(csr, reorder=not args.no_reorder)
    log(f"graph resident on device in {time.perf_counter() - t0:.1f}s (H2D + validation"
        f"{'' if args.no_reorder else ' + renumbering by degree'})")
    flags = _abi.GG_OUT_DEVICE
    tuned = None
    if args.options:
        fixed = json.loads(args.options)
        g.set_options(**fixed)
        tuned = {"fixed": fixed}
        log(f"options: {fixed}")
    elif not args.no_tune:
        t0 = time.perf_counter()
        best, trials = g.autotune(programs=_abi.GG_TUNE_BFS | _abi.GG_TUNE_SSSP, src=src, reps=3)
        defaults = gg.default_options().as_dict()
        tuned = {"seconds": round(time.perf_counter() - t0, 2), "trials": len(trials),
                 "baseline_ms": round(trials[0][3], 3), "tuned_ms": None,
                 "changed": {k: v for k, v in best.items() if v != defaults[k]}}
        log(f"autotune: {tuned}")

    def step(instrument):
        f = flags | (_abi.GG_INSTRUMENT if instrument else 0)
        _, sb = g.bfs(src, flags=f, out=out_bfs)
        _, ss = g.sssp(src, flags=f, out=out_sssp)
        return sb, ss

    for _ in range(args.warmup):
        step(False)
    # Timed region for `value`: no per-kernel events (they add a marker per
    # launch, which costs BFS about a quarter of its time).
    sb_list, ss_list = [], []
    g.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sb, ss = step(False)
        sb_list.append(sb)
        ss_list.append(ss)
    g.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # Second timed pass of the same steps with HIP events around every kernel
    # on the library's stream: the per-kernel durations behind `roofline`.
    g.profile_reset()
    g.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    g.synchronize()
    wall_instrumented = time.perf_counter() - t1
    prof = g.profile()
    launches = g.profile_launches()
    g.close()
    re_bfs = sb_list[-1]["reached_edges"]
    re_sssp = ss_list[-1]["reached_edges"]
    value = args.steps * (re_bfs + re_sssp) / wall / 1e9
    bfs_ms = float(np.median([s["gpu_ms"] for s in sb_list]))
    sssp_ms = float(np.median([s["gpu_ms"] for s in ss_list]))
    log(f"timed {args.steps} steps in {wall * 1e3:.1f} ms: {value:.2f} GTEPS "
        f"(bfs {bfs_ms:.3f} ms, sssp {sssp_ms:.3f} ms median)")
    if tuned is not None:
        tuned["tuned_ms"] = round(bfs_ms + sssp_ms, 3)
    name, ent = max(prof.items(), key=lambda kv: kv[1]["total_ms"])
    achieved = ent["alg_bytes"] / (ent["total_ms"] * 1e-3) / 1e9 if ent["total_ms"] > 0 else 0.0
    roofline = {
        "bound": "hbm", "kernel": name, "launches": ent["launches"],
        "avg_launch_us": ent["total_ms"] * 1e3 / max(ent["launches"], 1),
        "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
        "traffic": None, "traffic_source": None,
        "pass": f"second timed pass of {args.steps} steps with per-kernel HIP events "
                                   f"({wall_instrumented * 1e3:.1f} ms vs {wall * 1e3:.1f} ms without)",
    }
    tr = measured_traffic(name)
    if tr is not None:
        roofline["traffic"] = tr["bytes"]
        roofline["traffic_source"] = tr["source"]
    # Frontier-heavy launches: those expanding >= 1% of m edges (by each
    # kernel's per-edge byte model: push 12 B, top-down 8 B, bottom-up 4 B).
    per_edge = {"sssp.push": 12.0, "bfs.top_down": 8.0, "bfs.bottom_up": 4.0}
    heavy = {}
    for kname, bpe in per_edge.items():
        ls = [(ms, b) for nm, ms, b in launches if nm == kname and b >= 0.01 * csr.m * bpe and ms > 0]
        if ls:
            tb, tm = sum(b for _, b in ls), sum(ms for ms, _ in ls)
            best = max(b / (ms * 1e-3) / 1e9 for ms, b in ls)
            heavy[kname] = {"launches": len(ls), "achieved_GBps": round(tb / (tm * 1e-3) / 1e9, 1),
                            "frac": round(tb / (tm * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "best_launch_GBps": round(best, 1)}
    if name in heavy:
        roofline["frontier_heavy"] = heavy[name]
    kernels = {k: {"launches": v["launches"], "ms": round(v["total_ms"], 4),
                   "GBps": round(v["alg_bytes"] / (v["total_ms"] * 1e-3) / 1e9, 1) if v["total_ms"] > 0 else None}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["total_ms"])}
    per_alg = {
        "frontier_heavy": heavy,
        "bfs": {**gteps_row(re_bfs, sb_list[-1]["edges_inspected"], bfs_ms, True),
                "ms": bfs_ms, "levels": sb_list[-1]["iterations"], "pull_levels": sb_list[-1]["pull_iterations"],
                "edges_inspected": sb_list[-1]["edges_inspected"],
                "hbm_frac": sb_list[-1]["alg_bytes"] / (bfs_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "sssp": {**gteps_row(re_sssp, ss_list[-1]["edges_inspected"], sssp_ms, True),
                 "ms": sssp_ms, "iterations": ss_list[-1]["iterations"],
                 "edges_inspected": ss_list[-1]["edges_inspected"],
                 "hbm_frac": ss_list[-1]["alg_bytes"] / (sssp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    cpu = None
    if not args.no_cpu:
        import oracle

        t0 = time.perf_counter()
        exp_bfs = oracle.bfs(csr.row_offsets, csr.col_indices, src)
        tb = time.perf_counter() - t0
        t0 = time.perf_counter()
        exp_sssp = oracle.sssp(csr.row_offsets, csr.col_indices, csr.values, src)
        ts = time.perf_counter() - t0
        # Parity of the timed steps' last outputs (still on the device).
        per_alg["bfs"]["bit_exact"] = bool(np.array_equal(out_bfs.cpu().numpy(), exp_bfs))
        per_alg["sssp"]["bit_exact"] = bool(np.array_equal(out_sssp.cpu().numpy(), exp_sssp))
        cpu = {"value": (re_bfs + re_sssp) / (tb + ts) / 1e9, "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": f"one BFS ({tb:.2f} s) + one heap-Dijkstra SSSP ({ts:.2f} s) from the same source on "
                         f"the same R-MAT s{args.scale} graph, 1 thread"}
        log(f"cpu baseline: {cpu}")
        # The same BFS + SSSP on all host cores (OpenMP direction-optimising BFS,
        # bucketed SSSP with CAS-min; outputs equal to the serial oracle's).
        t0 = time.perf_counter()
        oracle.bfs_omp(csr.row_offsets, csr.col_indices, src)
        tbp = time.perf_counter() - t0
        t0 = time.perf_counter()
        oracle.sssp_omp(csr.row_offsets, csr.col_indices, csr.values, src)
        tsp = time.perf_counter() - t0
        cpu["parallel"] = {"value": (re_bfs + re_sssp) / (tbp + tsp) / 1e9, "unit": "GTEPS", "cores": oracle.threads(),
                           "sample": f"same BFS ({tbp:.3f} s) + SSSP ({tsp:.3f} s), OpenMP, {oracle.threads()} threads"}
        log(f"cpu baseline, all cores: {cpu['parallel']}")
    config = {"workload": f"rmat{args.scale}_undirected_bfs+sssp", "n": csr.n, "m_csr": csr.m, "edge_factor": 16,
              "weights": "int32 uniform [1,64]", "source": src, "step": "1 BFS + 1 SSSP",
              "teps_counting": "sum of CSR degree over reached vertices",
              "options": ("fixed " + json.dumps(tuned["fixed"]) if tuned and "fixed" in tuned
                          else "autotuned (gg_autotune, BFS + SSSP) before timing" if tuned else "defaults"),
              "vertex_order": "caller ids, stored renumbered by degree (GG_GRAPH_REORDER)" if not args.no_reorder
              else "as generated"}
    if tuned is not None and "fixed" not in tuned:
        per_alg["autotune"] = tuned
    # CC on the same graph (label propagation and hooking), outside the timed steps.
    if args.no_extras:
        return value, wall, roofline, kernels, per_alg, cpu, config
    with gg.Graph(csr, reorder=not args.no_reorder) as gc:
        per_alg["gas_path"] = gas_path_rows(gg, _abi, gc, csr, src, np)
        cc = {}
        exp_cc = None
        if not args.no_cpu:
            import oracle

            t0 = time.perf_counter()
            exp_cc, _ = oracle.cc(csr.row_offsets, csr.col_indices)
            cc["cpu_1core_ms"] = (time.perf_counter() - t0) * 1e3
            t0 = time.perf_counter()
            oracle.cc_omp(csr.row_offsets, csr.col_indices)
            cc["cpu_all_cores_ms"] = (time.perf_counter() - t0) * 1e3
            cc["cores"] = oracle.threads()
        for name, mode in (("labelprop", {}), ("hook", {"hook": True}), ("gas", {"gas": True})):
            ms, st = median_ms(lambda: gc.cc(flags=flags, out=out_bfs, **mode), 5)
            cc[name] = {"ms": ms, **gteps_row(st["reached_edges"], st["edges_inspected"], ms, True),
                        "iterations": st["iterations"]}
            if exp_cc is not None:
                cc[name]["bit_exact"] = bool(np.array_equal(out_bfs.cpu().numpy(), exp_cc))
        per_alg["cc"] = cc
    if not args.no_reorder:
        # The same BFS / SSSP on the graph stored in generated order.
        with gg.Graph(csr) as g0:
            b0, _ = median_ms(lambda: g0.bfs(src, flags=flags, out=out_bfs), 5)
            s0, _ = median_ms(lambda: g0.sssp(src, flags=flags, out=out_sssp), 5)
        per_alg["generated_order"] = {"bfs_ms": b0, "sssp_ms": s0}
    return value, wall, roofline, kernels, per_alg, cpu, config


def timed(fn, min_s=0.2, max_reps=200):
    """Seconds per call of a CPU function: repeated until min_s has passed."""
    t0 = time.perf_counter()
    reps = 0
    out = None
    while True:
        out = fn()
        reps += 1
        el = time.perf_counter() - t0
        if el >= min_s or reps >= max_reps:
            return el / reps, out


def gteps_row(reached_edges, edges_inspected, ms, symmetric):
    row = {"gteps": reached_edges / (ms * 1e-3) / 1e9, "work_gteps": edges_inspected / (ms * 1e-3) / 1e9}
    if symmetric:
        row["gteps_undirected"] = row["gteps"] / 2
    return row


def extra_config1(gg, workloads, np):
    """Config 1: BFS from vertex 0 on the Erdos-Renyi graph (n = 4096)."""
    import oracle

    c = workloads.config1_er()
    with gg.Graph(c) as g:
        ms, st = median_ms(lambda: g.bfs(0), 20)
        d, _ = g.bfs(0)
    t1, exp = timed(lambda: oracle.bfs(c.row_offsets, c.col_indices, 0))
    tp, _ = timed(lambda: oracle.bfs_omp(c.row_offsets, c.col_indices, 0))
    row = {"workload": f"erdos_renyi n={c.n} m_csr={c.m}, BFS from 0", "gpu_ms": ms,
           **gteps_row(st["reached_edges"], st["edges_inspected"], ms, True),
           "cpu_1core_ms": t1 * 1e3, "cpu_all_cores_ms": tp * 1e3, "cores": oracle.threads(),
           "speedup_vs_1core": t1 * 1e3 / ms, "speedup_vs_all_cores": tp * 1e3 / ms,
           "bit_exact": bool(np.array_equal(d, exp))}
    return row


def gas_path_rows(gg, _abi, g, csr, src, np):
    """The generic GAS program path (the reference's own schedules) on the
    headline graph: push-only BFS and GAS SSSP, with their byte models."""
    out = {}
    for name, fn in (("bfs_push_only", lambda: g.bfs(src, push_only=True)),
                     ("sssp_gas", lambda: g.sssp(src, flags=_abi.GG_SSSP_GAS))):
        ms, st = median_ms(fn, 5)
        out[name] = {"ms": ms, **gteps_row(st["reached_edges"], st["edges_inspected"], ms, True),
                     "iterations": st["iterations"], "hbm_frac": st["alg_bytes"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    return out


def extra_lattice(gg, workloads, np):
    """Config 3: CC on the 2048^2 percolated lattice, three GPU engines."""
    import oracle

    lat = workloads.config3_lattice(2048)
    out = {"workload": f"lattice 2048^2 p=0.6 + 0.1% long-range, n={lat.n} m_csr={lat.m}"}
    t0 = time.perf_counter()
    exp, sweeps = oracle.cc(lat.row_offsets, lat.col_indices)
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.cc_omp(lat.row_offsets, lat.col_indices)
    tp = time.perf_counter() - t0
    with gg.Graph(lat) as gl:
        for name, mode in (("labelprop", {}), ("labelprop_gas", {"gas": True}), ("hook", {"hook": True})):
            ms, st = median_ms(lambda: gl.cc(**mode), 3)
            lab, _ = gl.cc(**mode)
            out[name] = {"ms": ms, **gteps_row(st["reached_edges"], st["edges_inspected"], ms, True),
                         "iterations": st["iterations"], "bit_exact": bool(np.array_equal(lab, exp)),
                         "speedup_vs_1core": t1 * 1e3 / ms, "speedup_vs_all_cores": tp * 1e3 / ms}
    out.update({"cpu_1core_ms": t1 * 1e3, "cpu_1core_sweeps": sweeps, "cpu_all_cores_ms": tp * 1e3,
                "cores": oracle.threads()})
    return out


def extra_batch(gg, workloads, torch, np, scale, reorder):
    """Config 4: 128 SSSP + 64 PPR batched vs plain GPU launches vs the CPU."""
    import oracle

    c = workloads.config4_rmat(scale)
    src, seeds = workloads.config4_instances(c, 128, 64)
    out = {"graph": f"rmat{scale} directed, n={c.n}, m={c.m}", "vertex_order": "renumbered" if reorder else "generated"}
    with gg.Graph(c, directed=True, reorder=reorder) as g:
        g.batch(sssp_sources=src[:8], ppr_seeds=seeds[:4], tol_l1=1e-6)  # warm-up (allocations)
        g.sssp(int(src[0]))
        g.pagerank(teleport=np.eye(1, c.n, int(seeds[0]), dtype=np.float32)[0], tol_l1=1e-6)
        res = {}
        # Results stay on the device for both the batched and the plain runs.
        dist_dev = torch.empty((len(src), c.n), dtype=torch.int32, device="cuda")
        rank_dev = torch.empty((len(seeds), c.n), dtype=torch.float32, device="cuda")
        for label, s_list, p_list in (("sssp128", src, ()), ("ppr64", (), seeds), ("both", src, seeds)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, _, st = g.batch(sssp_sources=s_list, ppr_seeds=p_list, tol_l1=1e-6, flags=gg._abi.GG_OUT_DEVICE,
                               dist_out=dist_dev, rank_out=rank_dev)
            res[label] = (time.perf_counter() - t0) * 1e3
            if label == "sssp128":
                out["sssp128_union_edges"] = st["edges_inspected"]
                out["sssp128_iterations"] = st["iterations"]
            if label == "both":
                out["ppr_residual_l1"] = st["residual"]
        sample_s = [0, 37, 74, 111]
        sample_p = [0, 21, 42, 63]
        dist_rows = dist_dev[sample_s].cpu().numpy()
        rank_rows = rank_dev[sample_p].cpu().numpy()
        del dist_dev, rank_dev
        dev_i = torch.empty(c.n, dtype=torch.int32, device="cuda")
        dev_f = torch.empty(c.n, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in src:
            g.sssp(int(s), flags=gg._abi.GG_OUT_DEVICE, out=dev_i)
        plain_sssp = (time.perf_counter() - t0) * 1e3
        tele = torch.zeros(c.n, dtype=torch.float32, device="cuda")
        t0 = time.perf_counter()
        for s in seeds:
            tele.zero_()
            tele[int(s)] = 1.0
            torch.cuda.synchronize()  # torch's default stream -> the library's stream
            g.pagerank(teleport=tele, tol_l1=1e-6, flags=gg._abi.GG_OUT_DEVICE, out=dev_f)
        plain_ppr = (time.perf_counter() - t0) * 1e3
        # Batched BFS from the same 128 sources (two passes of 64) against 128
        # plain gg_bfs launches, results on the device in both.
        depth_dev = torch.empty((len(src), c.n), dtype=torch.int32, device="cuda")
        g.batch(bfs_sources=src[:64], flags=gg._abi.GG_OUT_DEVICE, depth_out=depth_dev)  # warm-up
        bfs_ms = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, _, _, bst = g.batch(bfs_sources=src, flags=gg._abi.GG_OUT_DEVICE, depth_out=depth_dev)
            bfs_ms.append((time.perf_counter() - t0) * 1e3)
        depth_rows = depth_dev[sample_s].cpu().numpy()
        del depth_dev
        reached = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s_ in src:
            _, pst = g.bfs(int(s_), flags=gg._abi.GG_OUT_DEVICE, out=dev_i)
            reached += pst["reached_edges"]
        plain_bfs = (time.perf_counter() - t0) * 1e3
        bb = float(np.median(bfs_ms))
        out["bfs128"] = {"batched_ms": bb, "plain_ms": plain_bfs, "speedup": plain_bfs / bb,
                         "batched_gteps": reached / (bb * 1e-3) / 1e9, "plain_gteps": reached / (plain_bfs * 1e-3) / 1e9,
                         "levels": bst["iterations"], "bottom_up_levels": bst["pull_iterations"],
                         "edges_inspected": bst["edges_inspected"]}
    out.update({"batched_ms": res, "plain_ms": {"sssp128": plain_sssp, "ppr64": plain_ppr,
                                                "both": plain_sssp + plain_ppr},
                "speedup": {"sssp128": plain_sssp / res["sssp128"], "ppr64": plain_ppr / res["ppr64"],
                            "both": (plain_sssp + plain_ppr) / res["both"]}})
    # Parity of sampled rows (SSSP bit-exact; PPR L1 against FP64 at 1e-14),
    # and the CPU's per-instance loop on the same sample, scaled to 128 + 64.
    co, ri, perm = oracle.csc(c.row_offsets, c.col_indices)
    ok, t1s, tps = True, 0.0, 0.0
    for k, i in enumerate(sample_s):
        t0 = time.perf_counter()
        exp = oracle.sssp(c.row_offsets, c.col_indices, c.values, int(src[i]))
        t1s += time.perf_counter() - t0
        t0 = time.perf_counter()
        oracle.sssp_omp(c.row_offsets, c.col_indices, c.values, int(src[i]))
        tps += time.perf_counter() - t0
        ok = ok and bool(np.array_equal(dist_rows[k], exp))
    errs, t1p, tpp = [], 0.0, 0.0
    for k, i in enumerate(sample_p):
        t = np.zeros(c.n)
        t[seeds[i]] = 1.0
        exp, _, _ = oracle.pagerank_f64(c.row_offsets, c.col_indices, t=t, tol=1e-14)
        errs.append(float(np.abs(rank_rows[k].astype(np.float64) - exp).sum()))
        t0 = time.perf_counter()
        oracle.pagerank_f64(c.row_offsets, c.col_indices, t=t, tol=1e-6)
        t1p += time.perf_counter() - t0
        t0 = time.perf_counter()
        oracle.pagerank_f64_omp(c.row_offsets, co, ri, t=t, tol=1e-6)
        tpp += time.perf_counter() - t0
    cpu1 = (t1s * 128 / len(sample_s) + t1p * 64 / len(sample_p)) * 1e3
    cpup = (tps * 128 / len(sample_s) + tpp * 64 / len(sample_p)) * 1e3
    bfs_ok = all(np.array_equal(depth_rows[k], oracle.bfs(c.row_offsets, c.col_indices, int(src[i])))
                 for k, i in enumerate(sample_s))
    out["parity"] = {"sssp_rows_bit_exact": ok, "sssp_rows_checked": len(sample_s), "bfs_rows_bit_exact": bfs_ok,
                     "ppr_l1_error_vs_fp64_max": max(errs), "ppr_rows_checked": len(sample_p),
                     "ppr_error_bound_at_tol": 0.85 / 0.15 * 1e-6}
    out["cpu"] = {"per_instance_loop_1core_ms": cpu1, "per_instance_loop_all_cores_ms": cpup,
                  "cores": oracle.threads(),
                  "sample": f"{len(sample_s)} of 128 SSSP sources (heap Dijkstra / OpenMP bucketed SSSP) and "
                            f"{len(sample_p)} of 64 PPR seeds (FP64 power iteration to L1 1e-6), scaled",
                  "speedup_batched_vs_1core": cpu1 / res["both"], "speedup_batched_vs_all_cores": cpup / res["both"]}
    return out


def extra_pagerank(gg, workloads, np, scale, reorder):
    """Config 5: FP32 PageRank and its adjoint, the FP64 oracle's L1 error and
    the CPU's FP64 power iteration beside them."""
    import oracle

    c = workloads.config5_rmat(scale)
    out = {"graph": f"rmat{scale} directed f32, n={c.n}, m={c.m}", "vertex_order": "renumbered" if reorder else "generated"}
    cvec = workloads.config5_loss_weights(c.n).astype(np.float32)
    with gg.Graph(c, directed=True, reorder=reorder) as g:
        g.pagerank(tol_l1=2e-7)
        ms, st = median_ms(lambda: g.pagerank(tol_l1=2e-7), 3)
        r, _ = g.pagerank(tol_l1=2e-7)
        out["forward"] = {"ms": ms, "sweeps": st["iterations"], "residual_l1": st["residual"],
                          "hbm_frac": st["alg_bytes"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          "work_gteps": st["edges_inspected"] / (ms * 1e-3) / 1e9}
        ams, ast = median_ms(lambda: g.pagerank_adjoint(r, cvec, tol_l1=2e-7)[1:], 3)
        out["adjoint"] = {"ms": ams, "sweeps": ast["iterations"],
                          "hbm_frac": ast["alg_bytes"] / (ams * 1e-3) / 1e9 / HBM_PEAK_GBS}
        # The same solves with XCD-sliced sweeps (opt-in: slower on this graph).
        g.set_options(xcd_slices=1)
        ms1, st1 = median_ms(lambda: g.pagerank(tol_l1=2e-7), 3)
        ams1, _ = median_ms(lambda: g.pagerank_adjoint(r, cvec, tol_l1=2e-7)[1:], 3)
        out["xcd_sliced"] = {"forward_ms": ms1, "forward_sweeps": st1["iterations"], "adjoint_ms": ams1}
    deg = c.degrees()
    rows = np.repeat(np.arange(c.n), deg)
    s = np.bincount(rows, weights=c.values.astype(np.float64), minlength=c.n)
    p = (c.values / np.repeat(np.where(deg > 0, s, 1.0), deg)).astype(np.float32).astype(np.float64)
    t0 = time.perf_counter()
    exp, it, _ = oracle.pagerank_f64(c.row_offsets, c.col_indices, p=p, tol=1e-12)
    out["forward"]["l1_error_vs_fp64"] = float(np.abs(r.astype(np.float64) - exp).sum())
    out["forward"]["fp64_oracle"] = {"tol_l1": 1e-12, "sweeps": it, "s": time.perf_counter() - t0}
    # CPU baseline: the same FP64 power iteration to the GPU's tolerance.
    t0 = time.perf_counter()
    _, it1, _ = oracle.pagerank_f64(c.row_offsets, c.col_indices, p=p, tol=2e-7)
    t1 = time.perf_counter() - t0
    co, ri, perm = oracle.csc(c.row_offsets, c.col_indices)
    pc = p[perm]
    t0 = time.perf_counter()
    oracle.pagerank_f64_omp(c.row_offsets, co, ri, p_csc=pc, tol=2e-7)
    tp = time.perf_counter() - t0
    out["cpu"] = {"forward_1core_ms": t1 * 1e3, "forward_all_cores_ms": tp * 1e3, "cores": oracle.threads(),
                  "sweeps": it1, "sample": "FP64 power iteration to L1 2e-7 on the same graph (push, 1 thread; "
                                          "pull over the CSC, OpenMP)",
                  "speedup_vs_1core": t1 * 1e3 / ms, "speedup_vs_all_cores": tp * 1e3 / ms}
    return out


def copy_probe(torch):
    """A plain device copy of known size, for calibrating FETCH_SIZE / WRITE_SIZE."""
    n = 256 << 20  # 1 GiB of float32
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": 4 * n}))


def lattice_gas_probe(gg, workloads, np):
    """Only the generic GAS CC program on the 2048^2 lattice (3 runs), for counter passes."""
    lat = workloads.config3_lattice(2048)
    with gg.Graph(lat) as gl:
        ref, _ = gl.cc()
        runs = []
        for _ in range(3):
            lab, st = gl.cc(gas=True)
            runs.append(st["gpu_ms"])
        print(json.dumps({"lattice_gas_cc_ms": runs, "iterations": st["iterations"],
                          "equal_to_labelprop": bool(np.array_equal(lab, ref))}))


def batch_bfs_probe(gg, workloads, torch, np):
    """Only the batched 128-source BFS on the R-MAT s20 directed graph (renumbered), 3 runs, for kernel traces."""
    c = workloads.config4_rmat(20)
    src, _ = workloads.config4_instances(c, 128, 64)
    with gg.Graph(c, directed=True, reorder=True) as g:
        depth = torch.empty((len(src), c.n), dtype=torch.int32, device="cuda")
        g.batch(bfs_sources=src[:64], flags=gg._abi.GG_OUT_DEVICE, depth_out=depth)  # warm-up
        runs = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, _, _, st = g.batch(bfs_sources=src, flags=gg._abi.GG_OUT_DEVICE, depth_out=depth)
            runs.append((time.perf_counter() - t0) * 1e3)
        print(json.dumps({"batched_bfs128_ms": runs, "levels": st["iterations"],
                          "bottom_up_levels": st["pull_iterations"], "edges_inspected": st["edges_inspected"]}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="headline workload only")
    ap.add_argument("--no-tune", action="store_true", help="time the default options (no gg_autotune)")
    ap.add_argument("--options", default=None,
                    help='fixed options as JSON, e.g. \'{"pull_alpha": 2}\' (no gg_autotune; for profiling runs)')
    ap.add_argument("--no-reorder", action="store_true", help="keep the R-MAT graphs in generated vertex order")
    ap.add_argument("--copy-probe", action="store_true", help="run only a 1 GiB device copy (PMC calibration)")
    ap.add_argument("--lattice-gas-probe", action="store_true",
                    help="run only the generic GAS CC on the 2048^2 lattice (for counter passes)")
    ap.add_argument("--batch-bfs-probe", action="store_true",
                    help="run only the batched 128-source BFS on the R-MAT s20 graph (for kernel traces)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_probe:
        copy_probe(torch)
        return

    import gpugraph_amd as gg
    from gpugraph_amd import _abi, workloads

    if args.batch_bfs_probe:
        batch_bfs_probe(gg, workloads, torch, np)
        return
    if args.lattice_gas_probe:
        lattice_gas_probe(gg, workloads, np)
        return

    value, wall, roofline, kernels, per_alg, cpu, config = headline(args, gg, _abi, workloads, torch, np)
    cpu_all = cpu.pop("parallel", None) if cpu else None
    cpu_main = cpu
    per_config = {}
    if not args.no_extras:
        per_config["1_er4096_bfs"] = extra_config1(gg, workloads, np)
        log(f"config 1: {per_config['1_er4096_bfs']}")
        per_config["2_rmat22_bfs_sssp_cc"] = {
            "bfs": per_alg["bfs"], "sssp": per_alg["sssp"], "cc": per_alg.get("cc"),
            "cpu_1core": cpu_main, "cpu_all_cores": cpu_all}
        per_config["3_lattice2048_cc"] = extra_lattice(gg, workloads, np)
        log(f"config 3: {per_config['3_lattice2048_cc']}")
        per_config["4_batched_rmat20"] = extra_batch(gg, workloads, torch, np, 20, not args.no_reorder)
        log(f"config 4: {per_config['4_batched_rmat20']}")
        per_config["5_pagerank_rmat22"] = extra_pagerank(gg, workloads, np, 22, not args.no_reorder)
        log(f"config 5: {per_config['5_pagerank_rmat22']}")
    line = {
        "metric": "GTEPS (BFS/SSSP/CC, R-MAT s22) + % HBM roofline; batched vs plain speedup",
        "value": value, "unit": "GTEPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic (seeded R-MAT generator)", "config": config,
        "roofline": roofline, "cpu_baseline": cpu_main, "cpu_baseline_all_cores": cpu_all, "per_algorithm": per_alg,
        "kernels": kernels, "per_config": per_config,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
