"""GG_GRAPH_REORDER: the graph is stored renumbered by degree, and every
entry point still takes and returns caller ids and edge order.  Results equal
the oracle's (BFS, SSSP, CC min-id labels, BFS parents: bit for bit;
PageRank and its adjoint: within the stated tolerances) and the unreordered
graph's."""
import numpy as np
import pytest

import gpugraph_amd as gg
import oracle
from gpugraph_amd import _abi, workloads

pytestmark = pytest.mark.gpu
L1_TOL = 1e-6


def l1(a, b):
    return float(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).sum())


def test_traversals_undirected(gpu_available):
    c = workloads.config2_rmat(scale=14)
    ro, col, w = c.row_offsets, c.col_indices, c.values
    srcs = (workloads.max_degree_vertex(c), 0, c.n - 1)
    with gg.Graph(c, reorder=True) as g:
        for src in srcs:
            exp_d = oracle.bfs(ro, col, src)
            for push in (False, True):
                d, _ = g.bfs(src, push_only=push)
                np.testing.assert_array_equal(d, exp_d)
            d, p, _ = g.bfs_tree(src)
            np.testing.assert_array_equal(d, exp_d)
            np.testing.assert_array_equal(p, oracle.bfs_parents(ro, col, exp_d, src))
            exp_s = oracle.sssp(ro, col, w, src)
            for flags in (0, _abi.GG_SSSP_GAS):
                s, _ = g.sssp(src, flags=flags)
                np.testing.assert_array_equal(s, exp_s)
        exp_cc = oracle.cc(ro, col)[0]
        for mode in ({}, {"gas": True}, {"hook": True}):
            lab, _ = g.cc(**mode)
            np.testing.assert_array_equal(lab, exp_cc)
        best, _ = g.autotune(programs=_abi.GG_TUNE_BFS | _abi.GG_TUNE_SSSP, src=srcs[0], reps=1)
        s, _ = g.sssp(srcs[0])
        np.testing.assert_array_equal(s, oracle.sssp(ro, col, w, srcs[0]))


def test_device_outputs_and_kats(gpu_available, kat_graphs):
    import torch

    for k in kat_graphs:
        c = gg.CSR(np.asarray(k["row_offsets"], np.int32), np.asarray(k["col_indices"], np.int32),
                   np.asarray(k["weights"], np.int32))
        with gg.Graph(c, directed=k["directed"], reorder=True) as g:
            for s, exp in k["bfs"].items():
                out = torch.full((c.n,), -5, dtype=torch.int32, device="cuda")
                g.bfs(int(s), flags=_abi.GG_OUT_DEVICE, out=out)
                g.synchronize()
                assert out.cpu().tolist() == exp, (k["name"], s)
            for s, exp in k["sssp"].items():
                d, _ = g.sssp(int(s))
                assert d.tolist() == exp, (k["name"], s)


def test_pagerank_adjoint_batch_and_update(gpu_available):
    c = gg.gen_rmat(13, 16, undirected=False, seed=81).with_weights("f32", symmetric=False, seed=82)
    deg = c.degrees()
    rows = np.repeat(np.arange(c.n), deg)
    ssum = np.bincount(rows, weights=c.values.astype(np.float64), minlength=c.n)
    p = (c.values / np.repeat(np.where(deg > 0, ssum, 1.0), deg)).astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(83)
    cv = rng.standard_normal(c.n)
    t = np.zeros(c.n)
    t[17] = 1.0
    with gg.Graph(c, directed=True, reorder=True) as g, gg.Graph(c, directed=True) as g0:
        r, _ = g.pagerank(0.85, tol_l1=2e-7)
        exp, _, _ = oracle.pagerank_f64(c.row_offsets, c.col_indices, p=p, tol=1e-14)
        assert l1(r, exp) <= L1_TOL
        rt, _ = g.pagerank(0.85, teleport=t.astype(np.float32), tol_l1=2e-7)
        assert l1(rt, oracle.pagerank_f64(c.row_offsets, c.col_indices, p=p, t=t, tol=1e-14)[0]) <= L1_TOL
        gw, gt, _ = g.pagerank_adjoint(r, cv.astype(np.float32), tol_l1=2e-7)
        r64, _, _ = oracle.pagerank_f64(c.row_offsets, c.col_indices, p=p, tol=1e-15)
        _, gw64, gt64, _ = oracle.pagerank_adjoint_f64(c.row_offsets, c.col_indices, c.values.astype(np.float64), r64,
                                                       cv, tol=1e-15)
        assert np.linalg.norm(gt - gt64) / np.linalg.norm(gt64) < 1e-4
        assert np.linalg.norm(gw - gw64) / np.linalg.norm(gw64) < 1e-4
        # the reference-form program (FP32 sums in a different order): same fixed point within its tolerance
        rr, _ = g.pagerank_reference()
        rr0, _ = g0.pagerank_reference()
        assert np.linalg.norm(rr.astype(np.float64) - rr0) / np.linalg.norm(rr0) < 0.02
        # edge values in caller order
        w2 = c.values.copy()
        w2[: c.m // 3] *= 4.0
        g.update_values(w2)
        g0.update_values(w2)
        np.testing.assert_allclose(g.pagerank(0.85, tol_l1=2e-7)[0], g0.pagerank(0.85, tol_l1=2e-7)[0], rtol=0,
                                   atol=1e-6)
    ci = gg.gen_rmat(13, 16, undirected=False, seed=84).with_weights("i32", 1, 64, symmetric=False, seed=85)
    cand = np.flatnonzero(ci.degrees() > 0)
    src = rng.choice(cand, 20, replace=False).astype(np.int32)
    seeds = rng.choice(cand, 6, replace=False).astype(np.int32)
    with gg.Graph(ci, directed=True, reorder=True) as g:
        dist, rank, _ = g.batch(sssp_sources=src, ppr_seeds=seeds, tol_l1=2e-7)
    for i in (0, 19):
        np.testing.assert_array_equal(dist[i], oracle.sssp(ci.row_offsets, ci.col_indices, ci.values, int(src[i])))
    for i in (0, 5):
        tt = np.zeros(ci.n)
        tt[seeds[i]] = 1.0
        assert l1(rank[i], oracle.pagerank_f64(ci.row_offsets, ci.col_indices, t=tt, tol=1e-14)[0]) <= L1_TOL


def test_vertex_order_is_by_degree(gpu_available):
    import ctypes as C

    c = gg.gen_rmat(12, 8, undirected=True, seed=86)
    with gg.Graph(c, reorder=True) as g:
        a, b = C.c_void_p(), C.c_void_p()
        _abi.check(g.lib.gg_graph_vertex_order(g.h, C.byref(a), C.byref(b)), "gg_graph_vertex_order")
        hip = C.CDLL("libamdhip64.so.7")
        hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        order = np.empty(c.n, np.int32)
        assert hip.hipMemcpy(order.ctypes.data, b.value, 4 * c.n, 2) == 0
    deg = c.degrees()
    np.testing.assert_array_equal(order, np.lexsort((np.arange(c.n), -deg)))
    with gg.Graph(c) as g:
        a, b = C.c_void_p(1), C.c_void_p(1)
        _abi.check(g.lib.gg_graph_vertex_order(g.h, C.byref(a), C.byref(b)), "gg_graph_vertex_order")
        assert a.value is None and b.value is None


def test_many_replays_host_and_device_outputs(gpu_available):
    """Runs longer than one captured replay (a lattice: BFS and SSSP take a
    few hundred levels / iterations) on a renumbered graph: the mapping to
    caller ids, enqueued with each replay's read-back, ends with the last
    replay's result, for host and device outputs, on first and repeated
    queries."""
    torch = pytest.importorskip("torch")
    lat = workloads.config3_lattice(256).with_weights("i32", 1, 64, symmetric=True, seed=7)
    ro, col, w = lat.row_offsets, lat.col_indices, lat.values
    src = workloads.max_degree_vertex(lat)
    exp_d = oracle.bfs(ro, col, src)
    exp_s = oracle.sssp(ro, col, w, src)
    assert exp_d.max() > 64  # many more levels than one replay holds
    with gg.Graph(lat, reorder=True) as g:
        dev = torch.device("cuda:0")
        out_d = torch.full((lat.n,), -7, dtype=torch.int32, device=dev)
        out_s = torch.full((lat.n,), -7, dtype=torch.int32, device=dev)
        for _ in range(2):
            d, _ = g.bfs(src)
            np.testing.assert_array_equal(d, exp_d)
            s, _ = g.sssp(src)
            np.testing.assert_array_equal(s, exp_s)
            g.bfs(src, flags=_abi.GG_OUT_DEVICE, out=out_d)
            g.sssp(src, flags=_abi.GG_OUT_DEVICE, out=out_s)
            np.testing.assert_array_equal(out_d.cpu().numpy(), exp_d)
            np.testing.assert_array_equal(out_s.cpu().numpy(), exp_s)
