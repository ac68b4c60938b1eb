"""Write tests/golden/kat_graphs.json: the reference's own small graphs as
known-answer tests (KATs).

Inputs (data files shipped with the reference, read here as text/binary data):
  smallRegressionGraphs/small.mtx         7 vertices, 11 weighted directed edges
  smallRegressionGraphs/trivialGraph.mtx  5 vertices, 5 edges (a chain + back edge)
  smallRegressionGraphs/chesapeake.mtx    39 vertices, pattern symmetric (mirrored)
  MPI/generator/Graphsizes + graphi_{0..3}_0
                                           32-vertex Graph500 sample split into 4
                                           2D partitions (int32 pairs, bfs.cu:579-606);
                                           merged, duplicates and self-loops kept

CSR construction follows b40c/graph/csr_graph.cuh:121-150 (stable sort by
(row, col), duplicates kept) and market.cuh:42-209 (1-based ids, mirror when
undirected).  Expected outputs are computed here by brute force that shares
no code with oracle/ or the library: pure-Python BFS, Bellman-Ford to a fixed
point, union-find for min-id components, and a dense FP64 linear solve for
PageRank.  Two hand-derived vectors from reading small.mtx and
trivialGraph.mtx are asserted before writing.

Run in a container that has the reference at /root/reference:
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import struct

import numpy as np

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat_graphs.json")
SSSP_INF = 100000000


def read_mtx(path, undirected):
    coo = []
    n = None
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("%"):
                continue
            parts = s.split()
            if n is None:
                n = int(parts[0])
                continue
            r, c = int(parts[0]) - 1, int(parts[1]) - 1
            v = float(parts[2]) if len(parts) > 2 else 1.0
            coo.append((r, c, v))
            if undirected:
                coo.append((c, r, v))
    return n, coo


def csr_from_coo(n, coo):
    coo = sorted(coo, key=lambda t: (t[0], t[1]))  # Python's sort is stable
    ro = [0] * (n + 1)
    for r, _, _ in coo:
        ro[r + 1] += 1
    for i in range(n):
        ro[i + 1] += ro[i]
    col = [c for _, c, _ in coo]
    val = [int(v) for _, _, v in coo]
    return ro, col, val


def bfs(n, ro, col, src):
    d = [-1] * n
    d[src] = 0
    q = [src]
    for u in q:
        for e in range(ro[u], ro[u + 1]):
            v = col[e]
            if d[v] == -1:
                d[v] = d[u] + 1
                q.append(v)
    return d


def bellman_ford(n, ro, col, w, src):
    d = [SSSP_INF] * n
    d[src] = 0
    changed = True
    while changed:
        changed = False
        for u in range(n):
            if d[u] >= SSSP_INF:
                continue
            for e in range(ro[u], ro[u + 1]):
                nd = d[u] + w[e]
                if nd < d[col[e]]:
                    d[col[e]] = nd
                    changed = True
    return d


def min_components(n, ro, col):
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for u in range(n):
        for e in range(ro[u], ro[u + 1]):
            a, b = find(u), find(col[e])
            if a != b:
                parent[max(a, b)] = min(a, b)
    return [find(v) for v in range(n)]


def pagerank_dense(n, ro, col, d=0.85):
    """Solve (I - d P^T - d/n 1 delta^T) r = (1-d)/n 1 exactly (FP64)."""
    A = np.eye(n)
    for u in range(n):
        deg = ro[u + 1] - ro[u]
        if deg == 0:
            A[:, u] -= d / n
        for e in range(ro[u], ro[u + 1]):
            A[col[e], u] -= d / deg
    return np.linalg.solve(A, np.full(n, (1 - d) / n)).tolist()


def symmetrize(n, ro, col):
    pairs = set()
    for u in range(n):
        for e in range(ro[u], ro[u + 1]):
            if u != col[e]:
                pairs.add((u, col[e]))
                pairs.add((col[e], u))
    sro, scol, _ = csr_from_coo(n, [(a, b, 1.0) for a, b in pairs])
    return sro, scol


def graph500_sample():
    with open(os.path.join(REF, "MPI/generator/Graphsizes"), "rb") as f:
        sizes = struct.unpack("<5i", f.read(20))
    coo = []
    for part in range(4):
        with open(os.path.join(REF, f"MPI/generator/graphi_{part}_0"), "rb") as f:
            data = struct.unpack(f"<{2 * sizes[part]}i", f.read(8 * sizes[part]))
        coo += [(data[2 * i], data[2 * i + 1], 1.0) for i in range(sizes[part])]
    return sizes[4], coo


def entry(name, n, coo, directed, note, synth_weights=False):
    ro, col, val = csr_from_coo(n, coo)
    # Deterministic int weights in [1, 64] for inputs that carry no values.
    if synth_weights:
        w = [1 + (3 * u + 7 * col[e]) % 64 for u in range(n) for e in range(ro[u], ro[u + 1])]
    else:
        w = val
    sro, scol = symmetrize(n, ro, col)
    g = dict(name=name, note=note, n=n, directed=directed, row_offsets=ro, col_indices=col, weights=w)
    g["bfs"] = {str(s): bfs(n, ro, col, s) for s in sorted({0, n // 2, n - 1})}
    g["sssp"] = {str(s): bellman_ford(n, ro, col, w, s) for s in sorted({0, n // 2, n - 1})}
    g["sym_row_offsets"], g["sym_col_indices"] = sro, scol
    g["cc_sym"] = min_components(n, sro, scol)
    g["pagerank_d085"] = pagerank_dense(n, ro, col)
    return g


def main():
    graphs = []
    n, coo = read_mtx(os.path.join(REF, "smallRegressionGraphs/small.mtx"), False)
    g = entry("small", n, coo, True, "smallRegressionGraphs/small.mtx, directed, weighted")
    assert g["bfs"]["0"] == [0, 1, 1, 1, 2, 2, 2], g["bfs"]["0"]      # read off the file by hand
    assert g["sssp"]["0"] == [0, 1, 4, 1, 2, 2, 2], g["sssp"]["0"]
    graphs.append(g)
    n, coo = read_mtx(os.path.join(REF, "smallRegressionGraphs/trivialGraph.mtx"), False)
    g = entry("trivial", n, coo, True, "smallRegressionGraphs/trivialGraph.mtx, directed")
    assert g["bfs"]["0"] == [0, 1, 2, 3, 4] and g["sssp"]["0"] == [0, 1, 2, 3, 4]
    graphs.append(g)
    n, coo = read_mtx(os.path.join(REF, "smallRegressionGraphs/chesapeake.mtx"), True)
    graphs.append(entry("chesapeake", n, coo, False, "chesapeake.mtx pattern symmetric, mirrored (undirected)",
                        synth_weights=True))
    n, coo = graph500_sample()
    graphs.append(entry("graph500_s5", n, coo, True,
                        "MPI/generator graphi_*_0 merged: 1024 pairs, duplicates and self-loops kept",
                        synth_weights=True))
    with open(OUT, "w") as f:
        json.dump({"graphs": graphs}, f, separators=(",", ":"))
    print(f"wrote {OUT}: " + ", ".join(f"{g['name']}(n={g['n']}, m={len(g['col_indices'])})" for g in graphs))


if __name__ == "__main__":
    main()
