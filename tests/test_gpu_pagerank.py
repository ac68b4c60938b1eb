"""GPU PageRank: FP32 compensated forward against the FP64 oracle (L1
tolerance 1e-6, stated per test), personalised teleports, the reference's own
PageRank program, the adjoint against the FP64 CPU adjoint and central finite
differences, and the torch.autograd wrapper."""
import numpy as np
import pytest

import gpugraph_amd as gg
import oracle
from gpugraph_amd import _abi, workloads

pytestmark = pytest.mark.gpu
L1_TOL = 1e-6  # FP32 forward vs FP64 oracle, summed over all vertices


def l1(a, b):
    return float(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).sum())


def normalised_p32(csr):
    """P_uv = w_uv / S_u in FP64, rounded to FP32 (the library's transition
    weights), returned as FP64 for the oracle."""
    deg = csr.degrees()
    w = csr.values.astype(np.float64)
    s = row_sums(csr, w)
    return (w / np.repeat(np.where(deg > 0, s, 1.0), deg)).astype(np.float32).astype(np.float64)


def row_sums(csr, w):
    rows = np.repeat(np.arange(csr.n), csr.degrees())
    return np.bincount(rows, weights=w, minlength=csr.n)


def test_kats(gpu_available, kat_graphs):
    for k in kat_graphs:
        csr = gg.CSR(k["row_offsets"], k["col_indices"])
        with gg.Graph(csr, directed=k["directed"]) as g:
            r, st = g.pagerank(0.85, tol_l1=2e-7)
            assert l1(r, k["pagerank_d085"]) <= L1_TOL, k["name"]
            assert st["residual"] <= 2e-7


@pytest.mark.parametrize("weighted", [False, True])
def test_rmat_directed(gpu_available, weighted):
    c = gg.gen_rmat(14, 16, undirected=False, seed=31)
    if weighted:
        c = c.with_weights("f32", symmetric=False, seed=32)
    p = normalised_p32(c) if weighted else None
    exp, _, _ = oracle.pagerank_f64(c.row_offsets, c.col_indices, p=p, tol=1e-14)
    with gg.Graph(c, directed=True) as g:
        r, st = g.pagerank(0.85, tol_l1=2e-7)
        r2, _ = g.pagerank(0.85, tol_l1=2e-7)
    assert l1(r, exp) <= L1_TOL
    np.testing.assert_array_equal(r, r2)  # reproducible bit for bit
    assert abs(float(np.sum(r, dtype=np.float64)) - 1.0) < 1e-5
    assert st["iterations"] > 5


def test_undirected_and_personalised(gpu_available):
    c = workloads.config1_er()
    rng = np.random.default_rng(33)
    seeds = rng.choice(c.n, 3, replace=False)
    with gg.Graph(c) as g:
        r, _ = g.pagerank(0.85, tol_l1=2e-7)
        assert l1(r, oracle.pagerank_f64(c.row_offsets, c.col_indices, tol=1e-14)[0]) <= L1_TOL
        for s in seeds:
            t = np.zeros(c.n)
            t[s] = 1.0
            r, _ = g.pagerank(0.85, teleport=t.astype(np.float32), tol_l1=2e-7)
            assert l1(r, oracle.pagerank_f64(c.row_offsets, c.col_indices, t=t, tol=1e-14)[0]) <= L1_TOL
            assert r[s] == r.max()


def test_reference_pagerank_program(gpu_available):
    """The reference's own PageRank (0.15 + 0.85 sum, per-vertex tolerance
    0.01) checked the way its driver checks it: relative L2 against CPUPR
    below tol (PageRank.cu:307-313, default tol 0.1); here also below 0.02."""
    c = gg.gen_rmat(13, 16, undirected=False, seed=34)
    exp, _ = oracle.pr_reference(c.row_offsets, c.col_indices)
    with gg.Graph(c, directed=True) as g:
        r, st = g.pagerank_reference()
    rel = np.linalg.norm(r.astype(np.float64) - exp) / np.linalg.norm(exp)
    assert rel < 0.02, rel
    assert st["iterations"] >= 2


@pytest.mark.parametrize("all_edges", [False, True])
def test_gas_gather_one_iteration(gpu_available, all_edges):
    """The generic engine's load-balanced frontier gather, pinned by one
    iteration of the reference PageRank program (iter_num = 1): every vertex
    starts at 0.15 in the frontier, so after one gather + apply
    r_v = 0.15 + 0.85 * sum_{u in N(v)} 0.15 / deg(u) exactly (N = in-edges,
    deg = out-degree; with GATHER_ALL_EDGES N = in- and out-edges, deg = out +
    in).  R-MAT scale 15 puts hub rows across many gather tiles; thresholds 0
    and 2^30 run the LBS and tiered expansions after it."""
    c = gg.gen_rmat(15, 16, undirected=False, seed=81)
    ro, col = c.row_offsets, c.col_indices
    n = c.n
    rows = np.repeat(np.arange(n), c.degrees())
    outdeg = c.degrees().astype(np.float64)
    indeg = np.bincount(col, minlength=n).astype(np.float64)
    deg = outdeg + indeg if all_edges else outdeg
    share = np.where(deg > 0, 0.15 / np.where(deg > 0, deg, 1.0), 0.0)
    acc = np.bincount(col, weights=share[rows], minlength=n)  # over in-edges
    if all_edges:
        acc += np.bincount(rows, weights=share[col], minlength=n)  # over out-edges
    exp = 0.15 + 0.85 * acc
    assert np.max(indeg) > 3 * 1024  # rows spanning several 1024-position tiles
    flags = _abi.GG_PR_ALL_EDGES if all_edges else 0
    with gg.Graph(c, directed=True) as g:
        for th in (0, 1 << 30):
            g.set_options(iter_num=1, threshold=th)
            r, st = g.pagerank_reference(flags=flags)
            assert st["iterations"] == 1
            np.testing.assert_allclose(r, exp, rtol=2e-6, atol=0)
        # Full runs through the device-driven and the host-driven (instrumented)
        # loops, each checked as the reference's driver checks the program.
        g.set_options(iter_num=(1 << 31) - 1, threshold=-1)
        r_dev, st_dev = g.pagerank_reference(flags=flags)
        r_host, st_host = g.pagerank_reference(flags=flags | _abi.GG_INSTRUMENT)
    assert st_dev["iterations"] > 1 and st_host["iterations"] > 1
    if not all_edges:
        ref, _ = oracle.pr_reference(ro, col)
        for r in (r_dev, r_host):
            assert np.linalg.norm(r - ref) / np.linalg.norm(ref) < 0.02


def test_adjoint_vs_fp64_and_finite_differences(gpu_available):
    c = gg.gen_rmat(12, 16, undirected=False, seed=35).with_weights("f32", symmetric=False, seed=36)
    n, m = c.n, c.m
    rng = np.random.default_rng(37)
    cv = rng.standard_normal(n)
    w64 = c.values.astype(np.float64)
    p = normalised_p32(c)
    with gg.Graph(c, directed=True) as g:
        r, _ = g.pagerank(0.85, tol_l1=2e-7)
        gw, gt, st = g.pagerank_adjoint(r, cv.astype(np.float32), tol_l1=2e-7)
    # FP64 CPU adjoint at the FP64 ranks (same transition weights).
    r64, _, _ = oracle.pagerank_f64(c.row_offsets, c.col_indices, p=p, tol=1e-15)
    lam, gw64, gt64, _ = oracle.pagerank_adjoint_f64(c.row_offsets, c.col_indices, w64, r64, cv, tol=1e-15)
    assert np.linalg.norm(gt - gt64) / np.linalg.norm(gt64) < 1e-4
    assert np.linalg.norm(gw - gw64) / np.linalg.norm(gw64) < 1e-4
    # Central differences on 256 random edges (FP64 forward, raw weights renormalised).
    deg = c.degrees()
    edges = rng.choice(m, 256, replace=False)

    def loss(wv):
        s = row_sums(c, wv)
        pv = wv / np.repeat(np.where(deg > 0, s, 1.0), deg)
        return float(cv @ oracle.pagerank_f64(c.row_offsets, c.col_indices, p=pv, tol=1e-15)[0])

    fd = np.empty(len(edges))
    for i, e in enumerate(edges):
        h = 1e-3 * max(abs(w64[e]), 1e-3)
        wp, wm = w64.copy(), w64.copy()
        wp[e] += h
        wm[e] -= h
        fd[i] = (loss(wp) - loss(wm)) / (2 * h)
    assert np.linalg.norm(gw[edges] - fd) / np.linalg.norm(fd) < 1e-4
    big = np.abs(fd) > 1e-2 * np.abs(fd).max()
    np.testing.assert_allclose(gw[edges][big], fd[big], rtol=1e-3)


def test_autograd_function(gpu_available):
    import torch

    from gpugraph_amd.autograd import pagerank

    c = gg.gen_rmat(11, 16, undirected=False, seed=38).with_weights("f32", symmetric=False, seed=39)
    rng = np.random.default_rng(40)
    cv = rng.standard_normal(c.n).astype(np.float32)
    with gg.Graph(c, directed=True) as g:
        w = torch.tensor(c.values, device="cuda", requires_grad=True)
        t = torch.full((c.n,), 1.0 / c.n, device="cuda", requires_grad=True)
        r = pagerank(g, w, t, tol_l1=2e-7, adj_tol=2e-7)
        loss = (torch.tensor(cv, device="cuda") * r).sum()
        loss.backward()
        torch.cuda.synchronize()
        r_ref, _ = g.pagerank(0.85, teleport=np.full(c.n, 1.0 / c.n, np.float32), tol_l1=2e-7)
        gw, gt, _ = g.pagerank_adjoint(r_ref, cv, tol_l1=2e-7)
    np.testing.assert_allclose(r.detach().cpu().numpy(), r_ref, rtol=0, atol=1e-9)
    np.testing.assert_allclose(w.grad.cpu().numpy(), gw, rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(t.grad.cpu().numpy(), gt, rtol=1e-5, atol=1e-9)


def test_autograd_two_forwards_one_backward(gpu_available):
    """Two forwards on one graph with different weights, then one backward
    through both: each call's gradient is its own adjoint (the second
    forward replaced the graph's resident weights)."""
    import torch

    from gpugraph_amd.autograd import pagerank

    c = gg.gen_rmat(11, 16, undirected=False, seed=43).with_weights("f32", symmetric=False, seed=44)
    rng = np.random.default_rng(45)
    c1 = rng.standard_normal(c.n).astype(np.float32)
    c2 = rng.standard_normal(c.n).astype(np.float32)
    w1v = c.values.copy()
    w2v = c.values.copy()
    w2v[: c.m // 2] *= 3.0
    with gg.Graph(c, directed=True) as g:
        w1 = torch.tensor(w1v, device="cuda", requires_grad=True)
        w2 = torch.tensor(w2v, device="cuda", requires_grad=True)
        r1 = pagerank(g, w1, tol_l1=2e-7, adj_tol=2e-7)
        r2 = pagerank(g, w2, tol_l1=2e-7, adj_tol=2e-7)
        loss = (torch.tensor(c1, device="cuda") * r1).sum() + (torch.tensor(c2, device="cuda") * r2).sum()
        loss.backward()
        torch.cuda.synchronize()
        expected = []
        for wv, cv in ((w1v, c1), (w2v, c2)):
            g.update_values(wv)
            rr, _ = g.pagerank(0.85, tol_l1=2e-7)
            expected.append(g.pagerank_adjoint(rr, cv, tol_l1=2e-7)[0])
    np.testing.assert_allclose(w1.grad.cpu().numpy(), expected[0], rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(w2.grad.cpu().numpy(), expected[1], rtol=1e-5, atol=1e-9)
    assert np.abs(expected[0] - expected[1]).max() > 1e-6  # the two calls really differ


@pytest.mark.parametrize("c_scale", [1e-3, 1e3])
def test_adjoint_convergence_status_scales_with_c(gpu_available, c_scale):
    """The adjoint's stopping test is residual <= tol_l1 * ||c||_1: scaling c
    scales the residual and the threshold alike, so the sweep count is the
    same, a run capped at that count returns GG_OK and one capped a sweep
    earlier returns GG_ERR_NOT_CONVERGED, whatever ||c||_1 is."""
    c = gg.gen_rmat(11, 16, undirected=False, seed=46).with_weights("f32", symmetric=False, seed=47)
    rng = np.random.default_rng(48)
    cv = rng.standard_normal(c.n)
    cv *= 1.0 / np.abs(cv).sum()  # ||c||_1 = 1
    with gg.Graph(c, directed=True) as g:
        r, _ = g.pagerank(0.85, tol_l1=2e-7)
        _, _, st1 = g.pagerank_adjoint(r, cv.astype(np.float32), tol_l1=1e-6)
        k = st1["iterations"]
        cs = (cv * c_scale).astype(np.float32)
        _, _, st = g.pagerank_adjoint(r, cs, tol_l1=1e-6)
        assert abs(st["iterations"] - k) <= 1
        k = st["iterations"]
        _, _, st = g.pagerank_adjoint(r, cs, tol_l1=1e-6, max_iter=k)
        assert st["residual"] <= 1e-6 * np.abs(cs.astype(np.float64)).sum()
        with pytest.raises(gg.GGError) as ei:
            g.pagerank_adjoint(r, cs, tol_l1=1e-6, max_iter=k - 1)
        assert ei.value.code == _abi.GG_ERR_NOT_CONVERGED


def test_update_values_changes_ranks(gpu_available):
    c = gg.gen_rmat(11, 8, undirected=False, seed=41).with_weights("f32", symmetric=False, seed=42)
    with gg.Graph(c, directed=True) as g:
        r1, _ = g.pagerank(0.85, tol_l1=2e-7)
        w2 = c.values.copy()
        w2[: c.m // 2] *= 5.0
        g.update_values(w2)
        r2, _ = g.pagerank(0.85, tol_l1=2e-7)
    c2 = gg.CSR(c.row_offsets, c.col_indices, w2)
    exp2, _, _ = oracle.pagerank_f64(c.row_offsets, c.col_indices, p=normalised_p32(c2), tol=1e-14)
    assert l1(r2, exp2) <= L1_TOL
    assert l1(r1, r2) > 1e-4


def test_errors_and_int_weights(gpu_available):
    c = workloads.config1_er()
    with gg.Graph(c.with_weights("i32", 1, 64, symmetric=True, seed=1)) as g:
        r, _ = g.pagerank(tol_l1=2e-7)  # int32 values are SSSP weights: PageRank is unweighted
        assert l1(r, oracle.pagerank_f64(c.row_offsets, c.col_indices, tol=1e-14)[0]) <= L1_TOL
    with gg.Graph(workloads.config1_er()) as g:
        with pytest.raises(gg.GGError) as e:
            g.pagerank(tol_l1=0.0, max_iter=3)
        assert e.value.code == _abi.GG_ERR_NOT_CONVERGED
        with pytest.raises(gg.GGError):
            g.pagerank(damping=1.0)


def star_ring(n):
    """Every vertex points to vertex 0 and to its ring successor: vertex 0 has
    n - 1 in-edges (more than 4096 per XCD slice for n > 32768)."""
    src = np.concatenate([np.arange(1, n), np.arange(n)])
    dst = np.concatenate([np.zeros(n - 1, np.int64), (np.arange(n) + 1) % n])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    ro = np.zeros(n + 1, np.int64)
    np.add.at(ro, src + 1, 1)
    return gg.CSR(np.cumsum(ro).astype(np.int32), dst.astype(np.int32))


@pytest.mark.parametrize("weighted", [False, True])
def test_xcd_sliced_sweeps(gpu_available, weighted):
    """Sliced sweeps (xcd_slices = 1: per-slice partials, short / wave / workgroup
    sub-rows, combine) against the FP64 oracle and the unsliced sweeps."""
    graphs = [gg.gen_rmat(14, 16, undirected=False, seed=43), star_ring(40000)]
    for c in graphs:
        if weighted:
            c = c.with_weights("f32", symmetric=False, seed=44)
        p = normalised_p32(c) if weighted else None
        exp, _, _ = oracle.pagerank_f64(c.row_offsets, c.col_indices, p=p, tol=1e-14)
        with gg.Graph(c, directed=True) as g:
            g.set_options(xcd_slices=1)
            r1, st1 = g.pagerank(0.85, tol_l1=2e-7)
            r1b, _ = g.pagerank(0.85, tol_l1=2e-7)
            g.set_options(xcd_slices=0)
            r0, _ = g.pagerank(0.85, tol_l1=2e-7)
        assert l1(r1, exp) <= L1_TOL and l1(r0, exp) <= L1_TOL
        np.testing.assert_array_equal(r1, r1b)
        assert st1["residual"] <= 2e-7


def test_xcd_sliced_adjoint_and_update(gpu_available):
    c = gg.gen_rmat(12, 16, undirected=False, seed=45).with_weights("f32", symmetric=False, seed=46)
    rng = np.random.default_rng(47)
    cv = rng.standard_normal(c.n)
    p = normalised_p32(c)
    r64, _, _ = oracle.pagerank_f64(c.row_offsets, c.col_indices, p=p, tol=1e-15)
    _, gw64, gt64, _ = oracle.pagerank_adjoint_f64(c.row_offsets, c.col_indices, c.values.astype(np.float64), r64, cv,
                                                   tol=1e-15)
    with gg.Graph(c, directed=True) as g:
        g.set_options(xcd_slices=1)
        r, _ = g.pagerank(0.85, tol_l1=2e-7)
        gw, gt, _ = g.pagerank_adjoint(r, cv.astype(np.float32), tol_l1=2e-7)
        assert np.linalg.norm(gt - gt64) / np.linalg.norm(gt64) < 1e-4
        assert np.linalg.norm(gw - gw64) / np.linalg.norm(gw64) < 1e-4
        # new weights reach the slice-ordered transition weights
        w2 = c.values.copy()
        w2[: c.m // 2] *= 5.0
        g.update_values(w2)
        r2, _ = g.pagerank(0.85, tol_l1=2e-7)
    exp2, _, _ = oracle.pagerank_f64(c.row_offsets, c.col_indices, p=normalised_p32(gg.CSR(c.row_offsets, c.col_indices, w2)),
                                     tol=1e-14)
    assert l1(r2, exp2) <= L1_TOL


def csr_from_edges(n, src, dst):
    src, dst = np.asarray(src, np.int64), np.asarray(dst, np.int64)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    ro = np.zeros(n + 1, np.int64)
    np.add.at(ro, src + 1, 1)
    return gg.CSR(np.cumsum(ro).astype(np.int32), dst.astype(np.int32))


def tile_shape_graphs():
    """Directed graphs whose in-rows (forward sweep) and out-rows (adjoint)
    cross the 1024-position merge-path tiles in every way: rows of 1023 / 1024
    / 1025 / 5000 edges spanning one to six tiles, mostly empty rows, and
    rows of one edge."""
    rng = np.random.default_rng(48)
    out = []
    for big in (1023, 1024, 1025, 5000):
        n = 3000
        src, dst = [], []
        for hub in (0, 7, 1500):  # three hubs: long in-rows and long out-rows
            s = rng.choice(n, big if big < n else n - 1, replace=big > n - 1)
            src += list(s)
            dst += [hub] * len(s)
            src += [hub] * len(s)
            dst += list(rng.integers(0, n, len(s)))
        ring = np.arange(0, n, 3)  # a third of the vertices on a ring, the rest with empty rows
        src += list(ring)
        dst += list(np.roll(ring, -1))
        out.append(csr_from_edges(n, src, dst))
    n = 20000  # 90 % empty in-rows, single edges elsewhere
    s = rng.integers(0, n, 2000)
    out.append(csr_from_edges(n, s, rng.choice(n, 2000, replace=False)))
    return out


@pytest.mark.parametrize("weighted", [False, True])
def test_segmented_sweep_tile_shapes(gpu_available, weighted):
    """Merge-path segmented sweeps: FP64 oracle agreement (L1 <= 1e-6) and the
    same bits for any grid size and run (the fold order is fixed by the
    tiles), forward and adjoint."""
    rng = np.random.default_rng(49)
    for c in tile_shape_graphs():
        if weighted:
            c = c.with_weights("f32", symmetric=False, seed=50)
        p = normalised_p32(c) if weighted else None
        exp, _, _ = oracle.pagerank_f64(c.row_offsets, c.col_indices, p=p, tol=1e-14)
        cv = rng.standard_normal(c.n)
        w64 = c.values.astype(np.float64) if weighted else None
        _, gw64, gt64, _ = oracle.pagerank_adjoint_f64(c.row_offsets, c.col_indices, w64, exp, cv, tol=1e-15)
        runs = []
        with gg.Graph(c, directed=True) as g:
            for grid in (64, 2048, 65536):
                g.set_options(expand_grid=grid)
                r, st = g.pagerank(0.85, tol_l1=2e-7)
                gw, gt, _ = g.pagerank_adjoint(r, cv.astype(np.float32), tol_l1=2e-7)
                runs.append((r, gw, gt))
        r, gw, gt = runs[0]
        assert l1(r, exp) <= L1_TOL
        assert st["residual"] <= 2e-7
        assert np.linalg.norm(gt - gt64) / np.linalg.norm(gt64) < 1e-4
        assert np.linalg.norm(gw - gw64) / np.linalg.norm(gw64) < 1e-4
        for r2, gw2, gt2 in runs[1:]:
            np.testing.assert_array_equal(r2, r)
            np.testing.assert_array_equal(gw2, gw)
            np.testing.assert_array_equal(gt2, gt)
