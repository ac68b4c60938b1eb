"""Batched instances (gg_batch_run): every SSSP instance bit-equal to the
oracle and to its plain single-source run; every personalised PageRank within
L1 1e-6 of the FP64 oracle with a one-hot teleport; passes beyond 128 / 64
instances; argument errors."""
import numpy as np
import pytest

import gpugraph_amd as gg
import oracle
from gpugraph_amd import workloads

pytestmark = pytest.mark.gpu


def l1(a, b):
    return float(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).sum())


@pytest.fixture(scope="module")
def rmat_directed():
    return workloads.config4_rmat(scale=13)


def test_sssp_batch_matches_oracle_and_plain(gpu_available, rmat_directed):
    c = rmat_directed
    src, _ = workloads.config4_instances(c, 130, 0)  # two passes: 128 + 2
    src[5] = src[4]  # a repeated source is two independent instances
    with gg.Graph(c, directed=True) as g:
        dist, _, st = g.batch(sssp_sources=src)
        for i in (0, 4, 5, 63, 64, 127, 128, 129):
            exp = oracle.sssp(c.row_offsets, c.col_indices, c.values, int(src[i]))
            np.testing.assert_array_equal(dist[i], exp, err_msg=f"instance {i}")
        for i in range(0, 130, 9):
            plain, _ = g.sssp(int(src[i]))
            np.testing.assert_array_equal(dist[i], plain)
    assert st["iterations"] >= 2


def test_ppr_batch(gpu_available, rmat_directed):
    c = rmat_directed
    _, seeds = workloads.config4_instances(c, 0, 64)
    with gg.Graph(c, directed=True) as g:
        _, rank, st = g.batch(ppr_seeds=seeds, damping=0.85, tol_l1=2e-7, max_iter=1000)
        for i in (0, 1, 31, 32, 63):
            t = np.zeros(c.n)
            t[seeds[i]] = 1.0
            exp, _, _ = oracle.pagerank_f64(c.row_offsets, c.col_indices, t=t, tol=1e-14)
            assert l1(rank[i], exp) <= 1e-6, i
            plain, _ = g.pagerank(0.85, teleport=t.astype(np.float32), tol_l1=2e-7)
            assert l1(rank[i], plain) <= 1e-6, i
    assert st["residual"] <= 2e-7


def test_mixed_batch_and_weighted_ppr(gpu_available):
    c = gg.gen_rmat(12, 16, undirected=False, seed=51)
    ci = c.with_weights("i32", 1, 64, symmetric=False, seed=52)
    rng = np.random.default_rng(53)
    cand = np.flatnonzero(c.degrees() > 0)
    src = rng.choice(cand, 40, replace=False).astype(np.int32)
    seeds = rng.choice(cand, 70, replace=False).astype(np.int32)  # two PPR passes: 64 + 6
    with gg.Graph(ci, directed=True) as g:
        dist, rank, _ = g.batch(sssp_sources=src, ppr_seeds=seeds, tol_l1=2e-7)
    for i in (0, 39):
        np.testing.assert_array_equal(dist[i], oracle.sssp(ci.row_offsets, ci.col_indices, ci.values, int(src[i])))
    for i in (0, 63, 64, 69):
        t = np.zeros(c.n)
        t[seeds[i]] = 1.0
        assert l1(rank[i], oracle.pagerank_f64(c.row_offsets, c.col_indices, t=t, tol=1e-14)[0]) <= 1e-6
    # F32-weighted transitions in the batched PPR
    cf = c.with_weights("f32", symmetric=False, seed=54)
    deg = cf.degrees()
    rows = np.repeat(np.arange(cf.n), deg)
    s = np.bincount(rows, weights=cf.values.astype(np.float64), minlength=cf.n)
    p = (cf.values / np.repeat(np.where(deg > 0, s, 1.0), deg)).astype(np.float32).astype(np.float64)
    with gg.Graph(cf, directed=True) as g:
        _, rank, _ = g.batch(ppr_seeds=seeds[:8], tol_l1=2e-7)
    for i in (0, 7):
        t = np.zeros(c.n)
        t[seeds[i]] = 1.0
        assert l1(rank[i], oracle.pagerank_f64(cf.row_offsets, cf.col_indices, p=p, t=t, tol=1e-14)[0]) <= 1e-6


def test_batch_device_outputs(gpu_available, rmat_directed):
    """GG_OUT_DEVICE: instance-major results written into device buffers equal
    the host-output run (n not a multiple of the 64-vertex transpose tile)."""
    import torch
    c = gg.gen_rmat(11, 8, undirected=False, seed=61).with_weights("i32", 1, 64, symmetric=False, seed=62)
    c = gg.CSR(*_trim(c, c.n - 5))
    src = np.array([0, 3, 17, c.n - 1], np.int32)
    seeds = np.array([2, 5, c.n - 2], np.int32)
    with gg.Graph(c, directed=True) as g:
        dh, rh, _ = g.batch(sssp_sources=src, ppr_seeds=seeds, tol_l1=2e-7)
        dd = torch.full((len(src), c.n), -7, dtype=torch.int32, device="cuda")
        rd = torch.full((len(seeds), c.n), -7.0, dtype=torch.float32, device="cuda")
        g.batch(sssp_sources=src, ppr_seeds=seeds, tol_l1=2e-7, flags=gg._abi.GG_OUT_DEVICE, dist_out=dd, rank_out=rd)
        g.synchronize()
    np.testing.assert_array_equal(dd.cpu().numpy(), dh)
    np.testing.assert_array_equal(rd.cpu().numpy(), rh)
    np.testing.assert_array_equal(dh[3], oracle.sssp(c.row_offsets, c.col_indices, c.values, int(src[3])))


def _trim(c, n):
    """Induced subgraph on vertices [0, n): (row_offsets, col_indices, values)."""
    keep = c.col_indices < n
    rows = np.repeat(np.arange(c.n), np.diff(c.row_offsets))
    sel = keep & (rows < n)
    deg = np.bincount(rows[sel], minlength=n)[:n]
    ro = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    return ro, c.col_indices[sel].astype(np.int32), c.values[sel]


def test_batch_errors(gpu_available):
    c = workloads.config1_er()
    with gg.Graph(c) as g:
        with pytest.raises(gg.GGError):
            g.batch(sssp_sources=[0])  # no int32 weights
        with pytest.raises(gg.GGError):
            g.batch(ppr_seeds=[c.n])  # seed out of range


def test_negative_weights_rejected(gpu_available):
    """A negative int32 weight makes shortest paths undefined on an undirected
    graph (a negative cycle); SSSP and batched SSSP refuse it with
    GG_ERR_INVALID instead of relaxing without end, also after values that
    were fine are replaced by negative ones."""
    from gpugraph_amd import _abi

    c = workloads.config1_er().with_weights("i32", 1, 64, symmetric=True, seed=54)
    bad = c.values.copy()
    k = int(np.flatnonzero(c.col_indices > 0)[0])
    bad[k] = -3
    with gg.Graph(c) as g:
        dist, _ = g.sssp(0)
        np.testing.assert_array_equal(dist, oracle.sssp(c.row_offsets, c.col_indices, c.values, 0))
        g.update_values(bad)
        for call in (lambda: g.sssp(0), lambda: g.sssp(0, flags=_abi.GG_SSSP_GAS), lambda: g.batch(sssp_sources=[0, 1])):
            with pytest.raises(gg.GGError) as ei:
                call()
            assert ei.value.code == _abi.GG_ERR_INVALID
        # PageRank-only batches do not read the weights as path lengths.
        g.batch(ppr_seeds=[0], tol_l1=1e-6)
        g.update_values(c.values)
        dist2, _ = g.sssp(0)
        np.testing.assert_array_equal(dist2, dist)


@pytest.mark.parametrize("undirected,reorder", [(False, False), (True, True), (False, True)])
def test_bfs_batch_matches_oracle(gpu_available, undirected, reorder):
    """Batched multi-source BFS: 70 sources (two passes, 64 + 6, one source
    repeated, one isolated vertex) bit-exact per instance against the oracle
    and against plain gg_bfs, host and device outputs, alone and beside a
    batched SSSP."""
    import torch

    c = gg.gen_rmat(13, 16, undirected=undirected, seed=55).with_weights("i32", 1, 64, symmetric=undirected,
                                                                         seed=56)
    rng = np.random.default_rng(57)
    src = rng.choice(c.n, 70, replace=False).astype(np.int32)
    src[3] = src[2]
    iso = np.flatnonzero(c.degrees() == 0)
    if len(iso):
        src[10] = iso[0]
    with gg.Graph(c, directed=not undirected, reorder=reorder) as g:
        _, _, depth, st = g.batch(bfs_sources=src)
        assert depth.shape == (70, c.n)
        for i in range(70):
            np.testing.assert_array_equal(depth[i], oracle.bfs(c.row_offsets, c.col_indices, int(src[i])),
                                          err_msg=f"instance {i}")
        for i in (0, 63, 64, 69):
            plain, _ = g.bfs(int(src[i]))
            np.testing.assert_array_equal(depth[i], plain)
        dev = torch.empty((70, c.n), dtype=torch.int32, device="cuda")
        g.batch(bfs_sources=src, flags=_abi_flags().GG_OUT_DEVICE, depth_out=dev)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(dev.cpu().numpy(), depth)
        dist, _, depth2, _ = g.batch(sssp_sources=src[:5], bfs_sources=src[:5])
        np.testing.assert_array_equal(depth2, depth[:5])
        for i in range(5):
            np.testing.assert_array_equal(dist[i], oracle.sssp(c.row_offsets, c.col_indices, c.values, int(src[i])))
    assert st["iterations"] >= 2


def test_bfs_batch_long_in_rows(gpu_available):
    """Batched BFS on a directed graph whose three hubs have in-rows of 3000
    to 5000 edges (several 1024-edge chunks each, scanned by different waves
    and folded with atomicOr), plus a chain and an isolated vertex: 64
    sources, every instance bit-exact against the oracle, with bottom-up
    levels taken."""
    rng = np.random.default_rng(71)
    n = 20000
    rows = [rng.integers(0, n - 20, 60000), np.repeat([5, 6, 7], 2000)]
    cols = [rng.integers(0, n - 20, 60000), rng.integers(0, n - 20, 6000)]
    for h, cnt in ((5, 5000), (6, 4000), (7, 3000)):  # long in-rows
        rows.append(rng.integers(0, n - 20, cnt))
        cols.append(np.full(cnt, h))
    chain = np.arange(n - 20, n - 2)
    rows += [chain, np.array([3])]
    cols += [chain + 1, np.array([n - 20])]
    row, col = np.concatenate(rows).astype(np.int64), np.concatenate(cols).astype(np.int64)
    keep = row != col
    key = np.unique(row[keep] * n + col[keep])
    row, col = key // n, key % n
    ro = np.zeros(n + 1, np.int64)
    np.add.at(ro, row + 1, 1)
    c = gg.CSR(np.cumsum(ro).astype(np.int32), col.astype(np.int32), None)
    indeg = np.bincount(col, minlength=n)
    assert indeg[5] > 3 * 1024 and indeg[7] > 2 * 1024
    src = rng.choice(n - 20, 64, replace=False).astype(np.int32)
    src[0] = n - 1  # isolated: no edge touches it
    with gg.Graph(c, directed=True) as g:
        _, _, depth, st = g.batch(bfs_sources=src)
        for i in range(64):
            np.testing.assert_array_equal(depth[i], oracle.bfs(c.row_offsets, c.col_indices, int(src[i])),
                                          err_msg=f"instance {i}")
    assert st["pull_iterations"] >= 1


def _abi_flags():
    from gpugraph_amd import _abi

    return _abi
