"""Benchmark: plain BFS + SSSP on R-MAT scale 22 (undirected, edge factor 16,
int32 weights in [1, 64], source = highest-degree vertex), one MI355X.

A step is one BFS and one SSSP from the source, graph resident in HBM and
outputs kept on the device.  value = GTEPS over the timed steps, counting
sum of CSR out-degrees of the reached vertices per traversal (the
reference's counting, Algorithms/BFS/bfs.cu:963-980); the undirected-edge
count is half of it.  Kernel timings come from HIP events recorded on the
library's stream inside the timed region (GG_INSTRUMENT); the roofline is
that of the dominant kernel class.  The CPU baseline is the oracle (a
single-threaded C restatement of the reference's CPUBFS and, for SSSP, a
heap Dijkstra with CPUSSSP's output) on the same graph and source.

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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="skip the lattice CC / ER lines")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU path is not built.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import gpugraph_amd as gg
    from gpugraph_amd import _abi, workloads

    t0 = time.perf_counter()
    csr = workloads.config2_rmat(args.scale)
    src = workloads.max_degree_vertex(csr)
    log(f"R-MAT s{args.scale}: n={csr.n} m={csr.m} src={src} (deg {csr.degrees()[src]}) "
        f"built in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda:0")
    out_bfs = torch.empty(csr.n, dtype=torch.int32, device=dev)
    out_sssp = torch.empty(csr.n, dtype=torch.int32, device=dev)
    t0 = time.perf_counter()
    g = gg.Graph(csr)
    log(f"graph resident on device in {time.perf_counter() - t0:.1f}s (H2D + validation)")
    flags = _abi.GG_OUT_DEVICE

    def step(instrument):
        f = flags | (_abi.GG_INSTRUMENT if instrument else 0)
        _, sb = g.bfs(src, flags=f, out=out_bfs.data_ptr())
        _, ss = g.sssp(src, flags=f, out=out_sssp.data_ptr())
        return sb, ss

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    g.profile_reset()
    sb_list, ss_list = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sb, ss = step(True)
        sb_list.append(sb)
        ss_list.append(ss)
    g.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = g.profile()
    re_bfs = sb_list[-1]["reached_edges"]
    re_sssp = ss_list[-1]["reached_edges"]
    total_edges = args.steps * (re_bfs + re_sssp)
    value = total_edges / wall / 1e9
    bfs_ms = float(np.median([s["gpu_ms"] for s in sb_list]))
    sssp_ms = float(np.median([s["gpu_ms"] for s in ss_list]))
    log(f"timed {args.steps} steps in {wall * 1e3:.1f} ms: {value:.2f} GTEPS "
        f"(bfs {bfs_ms:.3f} ms, sssp {sssp_ms:.3f} ms median)")

    # Dominant kernel class by event time over the timed region.
    name, ent = max(prof.items(), key=lambda kv: kv[1]["total_ms"])
    achieved = ent["alg_bytes"] / (ent["total_ms"] * 1e-3) / 1e9 if ent["total_ms"] > 0 else 0.0
    roofline = {
        "bound": "hbm", "kernel": name, "launches": ent["launches"],
        "avg_launch_us": ent["total_ms"] * 1e3 / max(ent["launches"], 1),
        "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
        "traffic": None,
    }
    kernels = {k: {"launches": v["launches"], "ms": round(v["total_ms"], 4),
                   "GBps": round(v["alg_bytes"] / (v["total_ms"] * 1e-3) / 1e9, 1) if v["total_ms"] > 0 else None}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["total_ms"])}

    per_alg = {
        "bfs": {"gteps": re_bfs / (bfs_ms * 1e-3) / 1e9, "gteps_undirected": re_bfs / 2 / (bfs_ms * 1e-3) / 1e9,
                "ms": bfs_ms, "levels": sb_list[-1]["iterations"], "pull_levels": sb_list[-1]["pull_iterations"],
                "edges_inspected": sb_list[-1]["edges_inspected"],
                "hbm_frac": sb_list[-1]["alg_bytes"] / (bfs_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "sssp": {"gteps": re_sssp / (sssp_ms * 1e-3) / 1e9, "ms": sssp_ms, "iterations": ss_list[-1]["iterations"],
                 "edges_inspected": ss_list[-1]["edges_inspected"],
                 "hbm_frac": ss_list[-1]["alg_bytes"] / (sssp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }

    extras = {}
    if not args.no_extras:
        lat = workloads.config3_lattice(2048)
        with gg.Graph(lat) as gl:
            for hook in (False, True):
                gl.cc(hook=hook)
                ts = []
                for _ in range(3):
                    _, st = gl.cc(hook=hook)
                    ts.append(st["gpu_ms"])
                ms = float(np.median(ts))
                extras["cc_lattice_" + ("hook" if hook else "labelprop")] = {
                    "ms": ms, "gteps": st["reached_edges"] / (ms * 1e-3) / 1e9, "iterations": st["iterations"]}
        log(f"lattice CC: {extras}")
    g.close()

    cpu = None
    if not args.no_cpu:
        import oracle

        t0 = time.perf_counter()
        oracle.bfs(csr.row_offsets, csr.col_indices, src)
        tb = time.perf_counter() - t0
        t0 = time.perf_counter()
        oracle.sssp(csr.row_offsets, csr.col_indices, csr.values, src)
        ts = time.perf_counter() - t0
        cpu = {"value": (re_bfs + re_sssp) / (tb + ts) / 1e9, "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": f"one BFS ({tb:.2f} s) + one heap-Dijkstra SSSP ({ts:.2f} s) from the same source on "
                         f"the same R-MAT s{args.scale} graph, 1 thread"}
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": "GTEPS (BFS/SSSP/CC, R-MAT s22) + % HBM roofline; batched vs plain speedup",
        "value": value, "unit": "GTEPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic (seeded R-MAT generator)",
        "config": {"workload": f"rmat{args.scale}_undirected_bfs+sssp", "n": csr.n, "m_csr": csr.m,
                   "edge_factor": 16, "weights": "int32 uniform [1,64]", "source": src,
                   "step": "1 BFS + 1 SSSP", "teps_counting": "sum of CSR degree over reached vertices"},
        "roofline": roofline, "cpu_baseline": cpu, "per_algorithm": per_alg, "kernels": kernels, "extras": extras,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
