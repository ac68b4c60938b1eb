"""Write tests/golden/pagerank_fd_rmat22.json: central finite differences of
the differentiable-PageRank loss on the full-size workload, for the GPU
adjoint's check at that size.

Workload: R-MAT scale 22 directed, edge factor 16, FP32 weights
(k + 1) / 2^24 (gpugraph_amd.workloads.config5_rmat), P_uv = w_uv / S_u in
FP64, dangling mass spread uniformly, uniform teleport, d = 0.85;
L = sum_v c_v r_v with c = workloads.config5_loss_weights (N(0, 1), seed 105).

For 256 CSR edges drawn with numpy's default_rng(205) (generator order, not
the library's renumbering), h_e = 1e-3 * max(|w_e|, 1e-3) and
fd_e = (L(w + h_e) - L(w - h_e)) / (2 h_e), each perturbed fixed point being
the FP64 fixed point at w (power iteration to L1 residual 1e-15) plus the
solution of the two fixed points' difference equation iterated to a relative
change of 1e-11 (oracle.pagerank_fd_f64, pinned against brute-force
differences of two full solves in tests/test_oracle.py).  The graph's digest
is stored so the test can confirm it regenerated the same graph.

Runs on the CPU (OpenMP; about ten minutes on 8 cores):
    python tests/golden/make_pagerank_fd.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pagerank_fd_rmat22.json")
NUM_EDGES = 256


def digest(c):
    x = hashlib.sha256()
    for a in (c.row_offsets, c.col_indices, c.values):
        x.update(np.ascontiguousarray(a).tobytes())
    return x.hexdigest()[:16]


def fixed_point(c):
    """FP64 ranks at the graph's weights (exact FP64 normalisation)."""
    import oracle

    deg = c.degrees()
    rows = np.repeat(np.arange(c.n), deg)
    w = c.values.astype(np.float64)
    s = np.bincount(rows, weights=w, minlength=c.n)
    p = w / np.repeat(np.where(deg > 0, s, 1.0), deg)
    r, sweeps, resid = oracle.pagerank_f64(c.row_offsets, c.col_indices, p=p, tol=1e-15, max_iter=100000)
    return r, sweeps, resid


def main(scale=22):
    import oracle
    from gpugraph_amd import workloads

    t0 = time.time()
    c = workloads.config5_rmat(scale)
    cv = workloads.config5_loss_weights(c.n)
    w = c.values.astype(np.float64)
    r, sweeps, resid = fixed_point(c)
    print(f"forward: {sweeps} sweeps, residual {resid:.2e}, {time.time() - t0:.1f} s", flush=True)
    rng = np.random.default_rng(205)
    edges = np.sort(rng.choice(c.m, NUM_EDGES, replace=False)).astype(np.int64)
    h = 1e-3 * np.maximum(np.abs(w[edges]), 1e-3)
    t1 = time.time()
    fd, fd_sweeps = oracle.pagerank_fd_f64(c.row_offsets, c.col_indices, w, r, cv, edges, h, tol_rel=1e-11)
    print(f"differences: {fd_sweeps} sweeps at most, {time.time() - t1:.1f} s", flush=True)
    doc = {
        "workload": f"R-MAT scale {scale} directed, edge factor 16, f32 weights (config5_rmat), d = 0.85, "
                    "uniform teleport, c = config5_loss_weights",
        "graph_digest": digest(c), "n": c.n, "m": c.m, "damping": 0.85,
        "forward": {"sweeps": int(sweeps), "residual_l1": float(resid)},
        "differences": {"tol_rel": 1e-11, "max_sweeps": int(fd_sweeps)},
        "edges": edges.tolist(), "h": h.tolist(), "fd": fd.tolist(),
    }
    with open(OUT, "w") as f:
        json.dump(doc, f)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
