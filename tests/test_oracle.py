"""Pins the CPU oracle against independent sources: the golden KATs (brute
force written separately), scipy.sparse.csgraph, dense FP64 linear solves,
finite differences and a hand-computed 3-vertex adjoint.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

import oracle

INF = 100000000


def random_csr(n, m, seed, symmetric=False, wmax=64):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    u, v = u[keep], v[keep]
    if symmetric:
        u, v = np.concatenate([u, v]), np.concatenate([v, u])
    key = np.unique(u.astype(np.int64) * n + v)
    u, v = (key // n).astype(np.int32), (key % n).astype(np.int32)
    ro = np.zeros(n + 1, np.int32)
    np.add.at(ro, u + 1, 1)
    ro = np.cumsum(ro).astype(np.int32)
    a, b = np.minimum(u, v).astype(np.int64), np.maximum(u, v).astype(np.int64)
    w = (1 + (a * 131 + b * 7919) % wmax).astype(np.int32)
    return ro, v, w


def to_scipy(ro, col, w=None):
    n = len(ro) - 1
    data = np.ones(len(col)) if w is None else np.asarray(w, np.float64)
    return sp.csr_matrix((data, col, ro), shape=(n, n))


def test_kats(kat_graphs):
    for g in kat_graphs:
        ro, col, w = g["row_offsets"], g["col_indices"], g["weights"]
        for s, exp in g["bfs"].items():
            assert oracle.bfs(ro, col, int(s)).tolist() == exp, g["name"]
        for s, exp in g["sssp"].items():
            assert oracle.sssp(ro, col, w, int(s), selection=True).tolist() == exp, g["name"]
            assert oracle.sssp(ro, col, w, int(s)).tolist() == exp, g["name"]
        lab, _ = oracle.cc(g["sym_row_offsets"], g["sym_col_indices"])
        assert lab.tolist() == g["cc_sym"].tolist(), g["name"]
        r, _, _ = oracle.pagerank_f64(ro, col, tol=1e-15)
        np.testing.assert_allclose(r, g["pagerank_d085"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_bfs_vs_scipy(seed):
    ro, col, _ = random_csr(3000, 9000, seed)
    for src in (0, 17, 2999):
        d = oracle.bfs(ro, col, src)
        ref = csg.shortest_path(to_scipy(ro, col), method="D", unweighted=True, indices=src)
        exp = np.where(np.isinf(ref), -1, ref).astype(np.int32)
        np.testing.assert_array_equal(d, exp)


@pytest.mark.parametrize("seed", [4, 5])
def test_sssp_vs_scipy(seed):
    ro, col, w = random_csr(2000, 8000, seed)
    for src in (0, 999):
        ref = csg.dijkstra(to_scipy(ro, col, w), indices=src)
        exp = np.where(np.isinf(ref), INF, ref).astype(np.int32)
        a = oracle.sssp(ro, col, w, src, selection=True)
        b = oracle.sssp(ro, col, w, src)
        np.testing.assert_array_equal(a, exp)
        np.testing.assert_array_equal(b, exp)


def test_sssp_unreached_and_sentinel():
    # Two components: vertices of the second keep the 100000000 sentinel.
    ro = np.array([0, 1, 1, 2, 2], np.int32)
    col = np.array([1, 3], np.int32)
    w = np.array([5, 7], np.int32)
    assert oracle.sssp(ro, col, w, 0).tolist() == [0, 5, INF, INF]
    assert oracle.sssp(ro, col, w, 0, selection=True).tolist() == [0, 5, INF, INF]


@pytest.mark.parametrize("seed", [6, 7])
def test_cc_vs_scipy(seed):
    ro, col, _ = random_csr(4000, 3000, seed, symmetric=True)
    lab, sweeps = oracle.cc(ro, col)
    ncomp, comp = csg.connected_components(to_scipy(ro, col), directed=False)
    # min vertex id per scipy component
    mins = np.full(ncomp, np.iinfo(np.int32).max)
    np.minimum.at(mins, comp, np.arange(len(comp)))
    np.testing.assert_array_equal(lab, mins[comp])
    assert sweeps >= 1


def dense_pr(ro, col, p, t, d):
    n = len(ro) - 1
    A = np.eye(n)
    for u in range(n):
        deg = ro[u + 1] - ro[u]
        if deg == 0:
            A[:, u] -= d / n
        for e in range(ro[u], ro[u + 1]):
            A[col[e], u] -= d * (p[e] if p is not None else 1.0 / deg)
    return np.linalg.solve(A, (1 - d) * t)


@pytest.mark.parametrize("weighted", [False, True])
def test_pagerank_f64_vs_dense_solve(weighted):
    ro, col, w = random_csr(300, 900, 8)
    n = len(ro) - 1
    rng = np.random.default_rng(9)
    p = None
    if weighted:
        raw = rng.uniform(0.01, 1.0, len(col))
        s = np.add.reduceat(raw, ro[:-1]) if len(col) else raw
        deg = np.diff(ro)
        srow = np.repeat(np.where(deg > 0, s, 1.0), deg)
        p = raw / srow
    t = rng.uniform(0, 1, n)
    t /= t.sum()
    r, it, res = oracle.pagerank_f64(ro, col, p=p, t=t, tol=1e-15)
    np.testing.assert_allclose(r, dense_pr(ro, col, p, t, 0.85), rtol=0, atol=1e-13)
    assert abs(r.sum() - 1.0) < 1e-12 and (r > 0).all()
    assert res <= 1e-15 and it > 10


def test_pr_reference_is_near_its_fixed_point():
    # CPUPR stops when no vertex moves by >= 0.01; the Gauss-Seidel iterate it
    # returns must lie close to the exact fixed point r = 0.15 + 0.85 P^T r.
    ro, col, _ = random_csr(500, 3000, 10)
    deg = np.diff(ro)
    r, iters = oracle.pr_reference(ro, col)
    n = len(ro) - 1
    A = np.eye(n)
    for u in range(n):
        for e in range(ro[u], ro[u + 1]):
            A[col[e], u] -= 0.85 / deg[u]
    exact = np.linalg.solve(A, np.full(n, 0.15))
    assert iters >= 2
    assert np.max(np.abs(r - exact)) < 0.1


def test_pr_reference_trivial_graph_hand_kat():
    """CPUPR (PageRank.cu:108-161) on smallRegressionGraphs/trivialGraph.mtx
    (0->1, 1->2, 2->3, 3->4, 4->3, all out-degree 1), worked by hand.
    Gauss-Seidel in vertex order from 0.15 everywhere: r0 = 0.15,
    r1 = 0.15 + 0.85 r0 = 0.2775, r2 = 0.385875 after the first sweep and
    fixed from then on.  r3_k = 0.15 + 0.85 (r2 + r4_{k-1}) with
    r4_{k-1} = 0.15 + 0.85 r3_{k-1} gives r3_k = a + q r3_{k-1}, a =
    0.60549375, q = 0.7225, r3_0 := 0 (r4_0 = 0.15), so r3_k = r3* (1 - q^k),
    r3* = a / (1 - q).  Sweep k >= 2 moves r3 by a q^(k-1) (r4 by 0.85 of
    that): >= 0.01 up to k = 13, 0.00885 at k = 14, so CPUPR stops after 14
    sweeps with r3 = r3* (1 - q^14) and r4 = 0.15 + 0.85 r3.  A Jacobi sweep,
    another threshold or a float 0.01f that changes the stop would change the
    count or the values."""
    ro = np.array([0, 1, 2, 3, 4, 5], np.int32)
    col = np.array([1, 2, 3, 4, 3], np.int32)
    a, q = 0.60549375, 0.7225
    r3 = a / (1 - q) * (1 - q ** 14)
    exp = np.array([0.15, 0.2775, 0.385875, r3, 0.15 + 0.85 * r3])
    r, sweeps = oracle.pr_reference(ro, col)
    assert sweeps == 14
    np.testing.assert_allclose(r, exp, rtol=0, atol=5e-6)


def test_pagerank_fd_matches_brute_force():
    """oracle.pagerank_fd_f64 (perturbed fixed points from the difference
    equation, batched) equals plain central differences of two full FP64
    solves per edge, on a graph small enough for the latter to carry no
    cancellation error; more than 32 edges so two sweep batches run."""
    ro, col, _ = random_csr(300, 1500, 13)
    n, m = len(ro) - 1, len(col)
    rng = np.random.default_rng(14)
    w = rng.uniform(0.05, 1.0, m)
    c = rng.standard_normal(n)
    deg = np.diff(ro)
    s = np.add.reduceat(w, ro[:-1])
    p = w / np.repeat(np.where(deg > 0, s, 1.0), deg)
    t = np.full(n, 1.0 / n)
    r, _, _ = oracle.pagerank_f64(ro, col, p=p, tol=1e-16, max_iter=100000)
    edges = rng.choice(m, 40, replace=False)
    h = 1e-3 * np.maximum(np.abs(w[edges]), 1e-3)
    fd, sweeps = oracle.pagerank_fd_f64(ro, col, w, r, c, edges, h, tol_rel=1e-13)
    brute = fd_grad(ro, col, w, t, c, 0.85, edges)
    assert np.linalg.norm(fd - brute) / np.linalg.norm(brute) < 1e-7
    assert sweeps > 5


def fd_grad(ro, col, w, t, c, d, edges, h_rel=1e-3):
    deg = np.diff(ro)

    def loss(wv):
        s = np.add.reduceat(wv, ro[:-1])
        p = wv / np.repeat(np.where(deg > 0, s, 1.0), deg)
        r, _, _ = oracle.pagerank_f64(ro, col, p=p, t=t, d=d, tol=1e-15)
        return float(c @ r)

    out = []
    for e in edges:
        h = h_rel * max(abs(w[e]), 1e-3)
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        out.append((loss(wp) - loss(wm)) / (2 * h))
    return np.array(out)


def test_adjoint_vs_finite_differences():
    ro, col, _ = random_csr(200, 800, 11)
    n, m = len(ro) - 1, len(col)
    rng = np.random.default_rng(12)
    w = rng.uniform(0.05, 1.0, m)
    t = np.full(n, 1.0 / n)
    c = rng.standard_normal(n)
    deg = np.diff(ro)
    s = np.add.reduceat(w, ro[:-1])
    p = w / np.repeat(np.where(deg > 0, s, 1.0), deg)
    r, _, _ = oracle.pagerank_f64(ro, col, p=p, t=t, tol=1e-15)
    lam, gw, gt, _ = oracle.pagerank_adjoint_f64(ro, col, w, r, c, tol=1e-15)
    edges = rng.choice(m, 24, replace=False)
    fd = fd_grad(ro, col, w, t, c, 0.85, edges)
    np.testing.assert_allclose(gw[edges], fd, rtol=1e-5, atol=1e-10)
    # dL/dt by finite differences on the teleport vector
    for v in (0, 5, 77):
        h = 1e-6
        tp, tm = t.copy(), t.copy()
        tp[v] += h
        tm[v] -= h
        lp = c @ oracle.pagerank_f64(ro, col, p=p, t=tp, tol=1e-15)[0]
        lm = c @ oracle.pagerank_f64(ro, col, p=p, t=tm, tol=1e-15)[0]
        assert abs(gt[v] - (lp - lm) / (2 * h)) < 1e-7


def test_adjoint_three_vertex_hand_case():
    # 0 -> 1, 0 -> 2, 1 -> 2, vertex 2 dangling; unweighted (P = 1/outdeg).
    # Adjoint: lambda = c + d (P lambda + delta (1.lambda)/n); with c = e_2 and
    # d = 0.5:  lambda_2 = 1 + (1/2)(S/3), lambda_1 = (1/2) lambda_2,
    # lambda_0 = (1/2)(lambda_1 + lambda_2)/2, S = sum(lambda).
    ro = np.array([0, 2, 3, 3], np.int32)
    col = np.array([1, 2, 2], np.int32)
    d = 0.5
    # Solve the 3x3 linear system by hand: unknowns l0, l1, l2.
    # l2 = 1 + (l0 + l1 + l2)/6 ; l1 = l2/2 ; l0 = (l1 + l2)/4 = 3 l2 / 8
    # l2 (1 - 1/6 (3/8 + 1/2 + 1)) = 1  ->  l2 = 1 / (1 - 15/48) = 48/33
    l2 = 48.0 / 33.0
    exp = np.array([3 * l2 / 8, l2 / 2, l2])
    r, _, _ = oracle.pagerank_f64(ro, col, d=d, tol=1e-15)
    lam, gw, gt, _ = oracle.pagerank_adjoint_f64(ro, col, None, r, np.array([0.0, 0.0, 1.0]), d=d, tol=1e-15)
    np.testing.assert_allclose(lam, exp, rtol=1e-12)
    np.testing.assert_allclose(gt, (1 - d) * exp, rtol=1e-12)
    # dL/dP_uv = d r_u lambda_v
    np.testing.assert_allclose(gw, [d * r[0] * exp[1], d * r[0] * exp[2], d * r[1] * exp[2]], rtol=1e-12)


def test_mutation_is_caught(kat_graphs):
    """A planted off-by-one in a copy of the BFS result must fail the KAT."""
    g = kat_graphs[0]
    d = oracle.bfs(g["row_offsets"], g["col_indices"], 0)
    d[3] += 1
    assert d.tolist() != g["bfs"]["0"]


def test_bfs_parents_kats(kat_graphs):
    """Hand-derived parents (smallest in-neighbour one level up)."""
    exp = {("small", 0): [0, 0, 0, 0, 1, 2, 3], ("small", 3): [-1, -1, -1, 3, -1, 3, 3],
           ("trivial", 0): [0, 0, 1, 2, 3], ("trivial", 4): [-1, -1, -1, 4, 4]}
    by_name = {g["name"]: g for g in kat_graphs}
    for (name, src), want in exp.items():
        g = by_name[name]
        d = oracle.bfs(g["row_offsets"], g["col_indices"], src)
        p = oracle.bfs_parents(g["row_offsets"], g["col_indices"], d, src)
        assert p.tolist() == want, (name, src)
        assert oracle.bfs_validate(g["row_offsets"], g["col_indices"], src, d, p) == 0


@pytest.mark.parametrize("seed", [6, 7])
def test_bfs_parents_and_validation_vs_scipy(seed):
    ro, col, _ = random_csr(3000, 9000, seed)
    rows = np.repeat(np.arange(3000), np.diff(ro))
    for src in (0, 1234):
        d = oracle.bfs(ro, col, src)
        # scipy's BFS tree (parent = discovering vertex) is another valid tree
        _, sp_pred = csg.breadth_first_order(to_scipy(ro, col), src, directed=True, return_predecessors=True)
        sp_pred = np.where(sp_pred < 0, -1, sp_pred).astype(np.int32)
        sp_pred[src] = src
        assert oracle.bfs_validate(ro, col, src, d, sp_pred) == 0
        p = oracle.bfs_parents(ro, col, d, src)
        assert oracle.bfs_validate(ro, col, src, d, p) == 0
        # brute force of the smallest-parent rule
        ok = (d[rows] >= 0) & (d[col] == d[rows] + 1)
        want = np.full(3000, -1, np.int64)
        for u, v in zip(rows[ok][::-1], col[ok][::-1]):  # ascending u wins: write descending, last write = min
            want[v] = u
        want[src] = src
        np.testing.assert_array_equal(p, want)


def test_bfs_validate_rejects_each_rule():
    ro, col, _ = random_csr(500, 2000, 8)
    src = 0
    d = oracle.bfs(ro, col, src)
    p = oracle.bfs_parents(ro, col, d, src)
    rows = np.repeat(np.arange(500), np.diff(ro))
    assert oracle.bfs_validate(ro, col, src, d, p) == 0
    reached = np.flatnonzero((d >= 2))
    v = int(reached[0])
    # 1: root not at depth 0
    d1 = d.copy(); d1[src] = 1
    assert oracle.bfs_validate(ro, col, src, d1, p) == 1
    # 2: a parent two levels up
    p2 = p.copy(); p2[v] = src
    assert oracle.bfs_validate(ro, col, src, d, p2) == 2
    # 2: two vertices of one level pointing at each other (a cycle)
    lvl = np.flatnonzero(d == d[v])
    a, b = int(lvl[0]), int(lvl[1])
    p3 = p.copy(); p3[a] = b; p3[b] = a
    assert oracle.bfs_validate(ro, col, src, d, p3) == 2
    # 3 and 4 on a hand graph: path 0->1->2->3, shortcut 0->3, isolated vertex 4
    hro = np.array([0, 2, 3, 4, 4, 4], np.int32)
    hcol = np.array([1, 3, 2, 3], np.int32)
    hd = oracle.bfs(hro, hcol, 0)
    hp = oracle.bfs_parents(hro, hcol, hd, 0)
    assert hd.tolist() == [0, 1, 2, 1, -1] and hp.tolist() == [0, 0, 1, 0, -1]
    assert oracle.bfs_validate(hro, hcol, 0, hd, hp) == 0
    # a consistent tree along the path, but the shortcut skips two levels
    assert oracle.bfs_validate(hro, hcol, 0, np.array([0, 1, 2, 3, -1]), np.array([0, 0, 1, 2, -1])) == 3
    # the isolated vertex given a parent
    assert oracle.bfs_validate(hro, hcol, 0, hd, np.array([0, 0, 1, 0, 0])) == 4
    # 5: a parent at the right level that is not an in-neighbour
    cand = [int(u) for u in np.flatnonzero(d == d[v] - 1) if u not in set(rows[col == v].tolist())]
    p6 = p.copy(); p6[v] = cand[0]
    assert oracle.bfs_validate(ro, col, src, d, p6) == 5


@pytest.mark.parametrize("seed", [9, 10])
def test_parallel_baselines_equal_serial(seed):
    """The OpenMP CPU baselines return exactly the serial oracle's outputs
    (directed graph with in-edges from oracle.csc; symmetric graph)."""
    ro, col, w = random_csr(5000, 40000, seed)
    co, ri = oracle.csc(ro, col)[:2]
    for src in (0, 4999):
        np.testing.assert_array_equal(oracle.bfs_omp(ro, col, src, co, ri), oracle.bfs(ro, col, src))
        for delta in (1, 16, 1000):
            np.testing.assert_array_equal(oracle.sssp_omp(ro, col, w, src, delta), oracle.sssp(ro, col, w, src))
    g = gg_rmat_sym(seed)
    src = int(np.argmax(np.diff(g[0])))
    np.testing.assert_array_equal(oracle.bfs_omp(g[0], g[1], src), oracle.bfs(g[0], g[1], src))
    np.testing.assert_array_equal(oracle.sssp_omp(g[0], g[1], g[2], src), oracle.sssp(g[0], g[1], g[2], src))
    assert oracle.threads() >= 1


def gg_rmat_sym(seed):
    import gpugraph_amd as gg

    c = gg.gen_rmat(13, 16, undirected=True, seed=seed).with_weights("i32", 1, 64, symmetric=True, seed=seed)
    return c.row_offsets, c.col_indices, c.values


@pytest.mark.parametrize("seed", [15, 16])
def test_parallel_cc_and_pagerank_equal_serial(seed):
    """The all-core CPU baselines compute the serial oracle's results: CC labels
    equal, PageRank within rounding (other summation order)."""
    ro, col, _ = random_csr(3000, 6000, seed, symmetric=True)
    lab, _ = oracle.cc(ro, col)
    lab2, _ = oracle.cc_omp(ro, col)
    np.testing.assert_array_equal(lab, lab2)
    ro, col, _ = random_csr(2000, 12000, seed)
    n = len(ro) - 1
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.05, 1.0, len(col))
    deg = np.diff(ro)
    s = np.add.reduceat(w, ro[:-1])
    p = w / np.repeat(np.where(deg > 0, s, 1.0), deg)
    t = rng.uniform(0, 1, n)
    t /= t.sum()
    co, ri, perm = oracle.csc(ro, col)
    for pp, tt in ((None, None), (p, t)):
        r, it, res = oracle.pagerank_f64(ro, col, p=pp, t=tt, tol=1e-13)
        r2, it2, res2 = oracle.pagerank_f64_omp(ro, co, ri, p_csc=None if pp is None else pp[perm], t=tt, tol=1e-13)
        assert abs(it - it2) <= 1
        assert np.abs(r - r2).sum() < 1e-12
