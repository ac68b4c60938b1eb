"""GPU parity of BFS / SSSP / CC through the C ABI: bit-exact against the
oracle on the golden KATs and seeded graphs, through both expand strategies
(LBS and tiered), and at full R-MAT scale-22 size through size-independent
characterisations of the results."""
import numpy as np
import pytest

import gpugraph_amd as gg
import oracle
from gpugraph_amd import _abi, workloads

pytestmark = pytest.mark.gpu
INF = gg.SSSP_INF


def csr_of(ro, col, w=None):
    return gg.CSR(np.asarray(ro, np.int32), np.asarray(col, np.int32), None if w is None else np.asarray(w, np.int32))


def test_kats(gpu_available, kat_graphs):
    for k in kat_graphs:
        with gg.Graph(csr_of(k["row_offsets"], k["col_indices"], k["weights"]), directed=k["directed"]) as g:
            for s, exp in k["bfs"].items():
                for push in (False, True):
                    d, _ = g.bfs(int(s), push_only=push)
                    assert d.tolist() == exp, (k["name"], s, push)
            for s, exp in k["sssp"].items():
                d, _ = g.sssp(int(s))
                assert d.tolist() == exp, (k["name"], s)
        with gg.Graph(csr_of(k["sym_row_offsets"], k["sym_col_indices"])) as g:
            for mode in ({}, {"gas": True}, {"hook": True}):
                lab, _ = g.cc(**mode)
                assert lab.tolist() == k["cc_sym"].tolist(), (k["name"], mode)


def test_config1_erdos_renyi(gpu_available):
    c = workloads.config1_er().with_weights("i32", 1, 64, symmetric=True, seed=101)
    with gg.Graph(c) as g:
        d, st = g.bfs(0)
        np.testing.assert_array_equal(d, oracle.bfs(c.row_offsets, c.col_indices, 0))
        assert st["reached_vertices"] == int((d >= 0).sum())
        assert st["reached_edges"] == int(c.degrees()[d >= 0].sum())
        dist, _ = g.sssp(0)
        np.testing.assert_array_equal(dist, oracle.sssp(c.row_offsets, c.col_indices, c.values, 0, selection=True))


@pytest.mark.parametrize("undirected", [True, False])
def test_rmat_bfs_sssp(gpu_available, undirected):
    c = gg.gen_rmat(15, 16, undirected=undirected, seed=21).with_weights("i32", 1, 64, symmetric=undirected,
                                                                         seed=22)
    deg = c.degrees()
    isolated = int(np.flatnonzero(deg == 0)[0])
    with gg.Graph(c, directed=not undirected) as g:
        for src in (workloads.max_degree_vertex(c), 1, isolated):
            exp = oracle.bfs(c.row_offsets, c.col_indices, src)
            for push in (False, True):
                d, st = g.bfs(src, push_only=push)
                np.testing.assert_array_equal(d, exp)
            exp_d = oracle.sssp(c.row_offsets, c.col_indices, c.values, src)
            for fl in (0, _abi.GG_SSSP_GAS):
                dist, _ = g.sssp(src, flags=fl)
                np.testing.assert_array_equal(dist, exp_d)


@pytest.mark.parametrize("delta", [1, 7, 32, 200, 2**31 - 1])
def test_sssp_near_far_bucket_widths(gpu_available, delta):
    c = gg.gen_rmat(14, 16, undirected=False, seed=26).with_weights("i32", 1, 64, symmetric=False, seed=27)
    src = workloads.max_degree_vertex(c)
    exp = oracle.sssp(c.row_offsets, c.col_indices, c.values, src)
    with gg.Graph(c, directed=True) as g:
        # pull_alpha 0: push only; 8: default switch; 2^20: pull whenever the frontier has edges
        # (and the previous iteration did not pull)
        for pull_alpha in (0, 8, 1 << 20):
            g.set_options(sssp_delta=delta, pull_alpha=pull_alpha)
            dist, st = g.sssp(src)
            np.testing.assert_array_equal(dist, exp)
            assert st["reached_vertices"] == int((exp < INF).sum())
            assert st["reached_edges"] == int(c.degrees()[exp < INF].sum())
            if pull_alpha == 1 << 20:
                assert st["pull_iterations"] >= 1


@pytest.mark.parametrize("threshold", [0, 1 << 30])
def test_expand_strategies_agree(gpu_available, threshold):
    """threshold 0 forces the LBS expand every iteration, 2^30 the tiered one."""
    c = gg.gen_rmat(14, 16, undirected=True, seed=23).with_weights("i32", 1, 64, symmetric=True, seed=24)
    src = workloads.max_degree_vertex(c)
    with gg.Graph(c) as g:
        for wave, block in ((32, 1024), (4, 16), (64, 1 << 20)):
            g.set_options(threshold=threshold, wave_degree=wave, block_degree=block)
            d, _ = g.bfs(src, push_only=True)
            np.testing.assert_array_equal(d, oracle.bfs(c.row_offsets, c.col_indices, src))
            dist, _ = g.sssp(src, flags=_abi.GG_SSSP_GAS)
            np.testing.assert_array_equal(dist, oracle.sssp(c.row_offsets, c.col_indices, c.values, src))
            lab, _ = g.cc(gas=True)
            np.testing.assert_array_equal(lab, oracle.cc(c.row_offsets, c.col_indices)[0])


def test_lattice_cc(gpu_available):
    """The three CC engines on the percolated lattice (many iterations of
    label propagation: several graph replays of the device-driven loop) and on
    R-MAT with its hubs (wave-cooperative rows); expand grid and wave split
    varied."""
    c = workloads.config3_lattice(side=256)
    exp, _ = oracle.cc(c.row_offsets, c.col_indices)
    with gg.Graph(c) as g:
        for mode in ({}, {"gas": True}, {"hook": True}):
            lab, st = g.cc(**mode)
            np.testing.assert_array_equal(lab, exp, err_msg=str(mode))
            assert st["reached_vertices"] == c.n
        lab, st = g.cc()
        assert st["iterations"] > 16  # more than one replay of the captured loop
        g.set_options(expand_grid=64, wave_degree=2)
        lab, _ = g.cc()
        np.testing.assert_array_equal(lab, exp)
    r = gg.gen_rmat(14, 16, undirected=True, seed=25)
    with gg.Graph(r) as g:
        exp_r = oracle.cc(r.row_offsets, r.col_indices)[0]
        for grid, wave in ((2048, 128), (64, 8)):
            g.set_options(expand_grid=grid, wave_degree=wave)
            lab, _ = g.cc()
            np.testing.assert_array_equal(lab, exp_r)


def test_direction_switch_parameters(gpu_available):
    c = gg.gen_rmat(16, 16, undirected=True, seed=25)
    src = workloads.max_degree_vertex(c)
    exp = oracle.bfs(c.row_offsets, c.col_indices, src)
    with gg.Graph(c) as g:
        for alpha, beta in ((1, 1), (14, 24), (1000000, 1)):
            g.set_options(bfs_alpha=alpha, bfs_beta=beta)
            d, st = g.bfs(src)
            np.testing.assert_array_equal(d, exp)
        g.set_options(bfs_alpha=1, bfs_beta=1000000)
        d, st = g.bfs(src)
        np.testing.assert_array_equal(d, exp)


@pytest.mark.parametrize("undirected", [True, False])
def test_bottom_up_every_level(gpu_available, undirected):
    """Bottom-up from the second level on (alpha 1, beta huge), bit-exact:
    an Erdos-Renyi graph whose vertex count is no multiple of 64 (partial
    last chunk and word), and an R-MAT of 2^20 vertices with unreachable ones,
    whose in-edges reach past the frontier bits kept in LDS (the first 2^19
    vertices).  From the hub and from vertex 1; device-driven and
    host-driven loops."""
    er = gg.gen_erdos_renyi(70001, 1200000, 77)
    rm = gg.gen_rmat(20, 8, undirected=undirected, seed=78)
    for c in (er, rm):
        hub = workloads.max_degree_vertex(c)
        with gg.Graph(c, directed=not undirected) as g:
            g.set_options(bfs_alpha=1, bfs_beta=1000000)
            for src in (hub, 1):
                exp = oracle.bfs(c.row_offsets, c.col_indices, src)
                for flags in (0, _abi.GG_INSTRUMENT):
                    d, st = g.bfs(src, flags=flags)
                    np.testing.assert_array_equal(d, exp, err_msg=str((c.n, src, flags)))
                    if src == hub:
                        assert st["pull_iterations"] >= 1, st


def test_device_outputs_and_instrumentation(gpu_available):
    import torch

    c = workloads.config1_er().with_weights("i32", 1, 64, symmetric=True, seed=101)
    out = torch.empty(c.n, dtype=torch.int32, device="cuda")
    with gg.Graph(c) as g:
        flags = _abi.GG_OUT_DEVICE | _abi.GG_INSTRUMENT
        g.bfs(0, flags=flags, out=out.data_ptr())
        g.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), oracle.bfs(c.row_offsets, c.col_indices, 0))
        g.sssp(0, flags=flags, out=out.data_ptr())
        g.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), oracle.sssp(c.row_offsets, c.col_indices, c.values, 0))
        g.cc(hook=True, flags=flags, out=out.data_ptr())
        g.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), oracle.cc(c.row_offsets, c.col_indices)[0])
        prof = g.profile()
        assert any(k.startswith("bfs.") for k in prof) and any(k.startswith("sssp.") for k in prof)
        assert all(v["launches"] > 0 and v["total_ms"] >= 0 for v in prof.values())


@pytest.mark.parametrize("offset", [0, 1, 3])
def test_sssp_collect_alignment_and_tail(gpu_available, offset):
    """The collect reads 16 dirty bytes per thread (the dirty array is padded
    to a multiple of 16) and, when the distance array is 16-byte aligned,
    16 distances as vectors; a caller's device array at another offset takes
    the one-load-per-vertex path.  n = 1000 + 13 is not a multiple of 16."""
    import torch

    c = gg.gen_erdos_renyi(1013, 1013 * 6, seed=51).with_weights("i32", 1, 30, seed=52)
    exp = {}
    buf = torch.empty(c.n + 8, dtype=torch.int32, device="cuda")
    out = buf[offset:offset + c.n]
    with gg.Graph(c, directed=True) as g:
        for opts in ({}, {"pull_alpha": 2, "sssp_delta": 8}):
            g.set_options(**gg.default_options().as_dict())
            g.set_options(**opts)
            for src in (0, c.n - 1, 517):
                if src not in exp:
                    exp[src] = oracle.sssp(c.row_offsets, c.col_indices, c.values, src)
                for fl in (_abi.GG_OUT_DEVICE, _abi.GG_OUT_DEVICE | _abi.GG_INSTRUMENT):
                    g.sssp(src, flags=fl, out=out.data_ptr())
                    g.synchronize()
                    np.testing.assert_array_equal(out.cpu().numpy(), exp[src], err_msg=str((offset, opts, src, fl)))
                    # the C ABI allows a NULL statistics pointer
                    out.fill_(-1)
                    _abi.check(g.lib.gg_sssp(g.h, src, out.data_ptr(), fl, None), "gg_sssp")
                    g.synchronize()
                    np.testing.assert_array_equal(out.cpu().numpy(), exp[src], err_msg=str((offset, opts, src, fl, "no stats")))


def test_sssp_replay_sizes_follow_previous_runs(gpu_available):
    """The device-driven SSSP sizes its first replay by the previous run on
    the same graph.  A short query, then one needing many more iterations
    than that (a 60-vertex unit-weight chain with unit buckets: about two
    loop iterations per vertex, several replays), then the short one with a
    replay longer than it needs: every result must match the oracle, and
    with pushes only the iteration counts must match the host-driven loop."""
    r = gg.gen_rmat(12, 8, undirected=True, seed=41).with_weights("i32", 1, 64, symmetric=True, seed=42)
    L = 60
    n = r.n + L
    rows = [np.repeat(np.arange(r.n), np.diff(r.row_offsets))]
    cols = [r.col_indices.astype(np.int64)]
    vals = [r.values.astype(np.int64)]
    a = np.arange(r.n, r.n + L - 1)
    rows += [a, a + 1]
    cols += [a + 1, a]
    vals += [np.ones(L - 1, np.int64)] * 2
    row, col, val = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    order = np.lexsort((col, row))
    row, col, val = row[order], col[order], val[order]
    ro = np.zeros(n + 1, np.int64)
    np.add.at(ro, row + 1, 1)
    c = csr_of(np.cumsum(ro), col, val)
    hub = workloads.max_degree_vertex(c)
    queries = (hub, r.n, hub, r.n + L // 2)
    with gg.Graph(c) as g:
        for opts in ({"pull_alpha": 2, "sssp_delta": 1}, {"pull_alpha": 0, "sssp_delta": 1}):
            g.set_options(**gg.default_options().as_dict())
            g.set_options(**opts)
            for src in queries:
                exp = oracle.sssp(c.row_offsets, c.col_indices, c.values, int(src))
                d, st = g.sssp(int(src))
                np.testing.assert_array_equal(d, exp, err_msg=str((opts, src)))
                if opts["pull_alpha"] == 0:
                    d2, st2 = g.sssp(int(src), flags=_abi.GG_INSTRUMENT)
                    np.testing.assert_array_equal(d2, exp)
                    assert st["iterations"] == st2["iterations"], (src, st["iterations"], st2["iterations"])
                    if src == r.n:
                        assert st["iterations"] > 40  # the chain query ran past several replays


@pytest.mark.parametrize("undirected", [True, False])
def test_device_driven_loops_match_host_loops(gpu_available, undirected):
    """BFS and SSSP run their iteration loops on the device (captured
    graphs) unless GG_INSTRUMENT asks for per-kernel events, which uses the
    host-driven loops.  Both must give the same results and statistics,
    through option changes and output buffers that force a new capture, and
    under an iteration cap."""
    import torch

    c = gg.gen_rmat(15, 16, undirected=undirected, seed=26).with_weights("i32", 1, 64, symmetric=undirected, seed=27)
    src = workloads.max_degree_vertex(c)
    exp_b = oracle.bfs(c.row_offsets, c.col_indices, src)
    exp_s = oracle.sssp(c.row_offsets, c.col_indices, c.values, src)
    # BFS statistics are fixed by the graph.  SSSP's are not: a pull reads
    # distances other workgroups lower during the same pass, and a push can
    # lower the distance of a vertex of its own frontier, whose edges then go
    # out with the old or the new value depending on which workgroup ran
    # first (the vertex is expanded again in the first case).  The results,
    # and the reached counts, are fixed.
    keys = {"bfs": ("iterations", "pull_iterations", "reached_vertices", "reached_edges", "edges_inspected"),
            "sssp": ("reached_vertices", "reached_edges")}
    outs = [torch.empty(c.n, dtype=torch.int32, device="cuda") for _ in range(2)]
    with gg.Graph(c, directed=not undirected, reorder=undirected) as g:
        for opts in ({}, {"pull_alpha": 2, "sssp_delta": 64}, {"bfs_alpha": 1, "bfs_beta": 1000000, "lbs_tile": 512},
                     {"pull_alpha": 1, "expand_grid": 64}):
            g.set_options(**gg.default_options().as_dict())
            g.set_options(**opts)
            for out in outs:
                for prog, exp in (("bfs", exp_b), ("sssp", exp_s)):
                    stats = {}
                    for fl in (_abi.GG_OUT_DEVICE, _abi.GG_OUT_DEVICE | _abi.GG_INSTRUMENT):
                        _, st = getattr(g, prog)(src, flags=fl, out=out.data_ptr())
                        g.synchronize()
                        np.testing.assert_array_equal(out.cpu().numpy(), exp, err_msg=str((prog, opts, fl)))
                        stats[fl] = {k: st[k] for k in keys[prog]}
                    assert stats[_abi.GG_OUT_DEVICE] == stats[_abi.GG_OUT_DEVICE | _abi.GG_INSTRUMENT], (prog, opts)
        # iteration cap: both loops stop after the same number of levels / iterations
        g.set_options(**gg.default_options().as_dict())
        g.set_options(iter_num=2)
        for prog in ("bfs", "sssp"):
            res = []
            for fl in (0, _abi.GG_INSTRUMENT):
                d, st = getattr(g, prog)(src, flags=fl)
                res.append((d, st["iterations"]))
            np.testing.assert_array_equal(res[0][0], res[1][0])
            assert res[0][1] == res[1][1] == 2, prog


@pytest.mark.parametrize("hi", [64, 1 << 24])
def test_sssp_pull_packed_and_wide_weights(gpu_available, hi):
    """Pull steps read packed in-edges (source and weight in one word) when
    ids and weights fit 32 bits together (weights up to 64 on 2^14 vertices),
    and the two-stream in-edges otherwise (weights up to 2^24); both against
    the oracle, with and without renumbering."""
    for undirected in (True, False):
        c = gg.gen_rmat(14, 16, undirected=undirected, seed=28).with_weights("i32", 1, hi, symmetric=undirected, seed=29)
        src = workloads.max_degree_vertex(c)
        exp = oracle.sssp(c.row_offsets, c.col_indices, c.values, src)
        for reorder in (False, True):
            pulls = 0
            with gg.Graph(c, directed=not undirected, reorder=reorder) as g:
                for alpha in (2, 4, 16):  # pull when the frontier's out-edges exceed m / alpha
                    # buckets as wide as a few edges, so frontiers grow large enough to pull
                    g.set_options(pull_alpha=alpha, sssp_delta=4 * hi)
                    for fl in (0, _abi.GG_INSTRUMENT):
                        d, st = g.sssp(src, flags=fl)
                        np.testing.assert_array_equal(d, exp, err_msg=str((hi, undirected, reorder, alpha, fl)))
                        pulls += st["pull_iterations"]
            assert pulls > 0


def test_error_paths(gpu_available):
    c = workloads.config1_er()
    with gg.Graph(c) as g:
        with pytest.raises(gg.GGError) as e:
            g.bfs(c.n)
        assert e.value.code == _abi.GG_ERR_INVALID
        with pytest.raises(gg.GGError) as e:
            g.sssp(0)  # no int32 weights on this graph
        assert e.value.code == _abi.GG_ERR_INVALID
        with pytest.raises(gg.GGError):
            g.set_options(wave_degree=0)
    with gg.Graph(c, directed=True) as g:
        for hook in (False, True):
            with pytest.raises(gg.GGError) as e:
                g.cc(hook=hook)
            assert e.value.code == _abi.GG_ERR_UNSUPPORTED
    # An undirected graph must be symmetric: 0->1 without 1->0 is rejected.
    with pytest.raises(gg.GGError) as e:
        gg.Graph(csr_of([0, 1, 1], [1]))
    assert e.value.code == _abi.GG_ERR_INVALID


def test_undirected_asymmetric_weights_and_unsorted_rows(gpu_available):
    """Mirrored entries with different weights: the pull step must read the
    in-edge's own weight.  Unsorted rows: symmetry unverifiable, push only."""
    c = gg.gen_rmat(13, 16, undirected=True, seed=28)
    rows = np.repeat(np.arange(c.n), c.degrees())
    w = (1 + (rows * 7 + c.col_indices * 3) % 64).astype(np.int32)  # w(u,v) != w(v,u)
    src = workloads.max_degree_vertex(c)
    exp = oracle.sssp(c.row_offsets, c.col_indices, w, src)
    with gg.Graph(gg.CSR(c.row_offsets, c.col_indices, w)) as g:
        g.set_options(pull_alpha=1 << 20)
        dist, st = g.sssp(src)
        np.testing.assert_array_equal(dist, exp)
        assert st["pull_iterations"] >= 1
    perm_col = c.col_indices.copy()
    perm_w = w.copy()
    for u in range(c.n):  # reverse every row
        b, e = c.row_offsets[u], c.row_offsets[u + 1]
        perm_col[b:e] = perm_col[b:e][::-1]
        perm_w[b:e] = perm_w[b:e][::-1]
    with gg.Graph(gg.CSR(c.row_offsets, perm_col, perm_w)) as g:
        g.set_options(pull_alpha=1 << 20)
        dist, st = g.sssp(src)
        np.testing.assert_array_equal(dist, exp)
        d, _ = g.bfs(src)
        np.testing.assert_array_equal(d, oracle.bfs(c.row_offsets, c.col_indices, src))


@pytest.fixture(scope="module")
def rmat22():
    return workloads.config2_rmat(22)


def test_config2_full_size(gpu_available, rmat22):
    """R-MAT scale 22 (4.19M vertices, 128M CSR entries): BFS bit-exact
    against the oracle; SSSP checked by the exact characterisation of
    shortest-path distances with positive weights (no edge can relax them,
    and every reached vertex except the source has a tight in-edge)."""
    c = rmat22
    src = workloads.max_degree_vertex(c)
    ro, col, w = c.row_offsets, c.col_indices, c.values
    rows = np.repeat(np.arange(c.n, dtype=np.int64), c.degrees())
    with gg.Graph(c) as g:
        d, st = g.bfs(src)
        np.testing.assert_array_equal(d, oracle.bfs(ro, col, src))
        assert st["pull_iterations"] >= 1
        dist, st = g.sssp(src)
    assert dist[src] == 0
    du = dist[rows].astype(np.int64)
    dv = dist[col].astype(np.int64)
    reached_u = du < INF
    assert np.all(dv[reached_u] <= du[reached_u] + w[reached_u])  # nothing relaxes
    tight = np.zeros(c.n, bool)
    hit = reached_u & (dv == du + w)
    tight[col[hit]] = True
    reached = dist < INF
    reached[src] = False
    assert np.all(tight[reached])  # every reached vertex has a shortest-path parent
    assert np.array_equal(dist < INF, d >= 0)  # same reachable set as BFS (undirected)


def test_bfs_tree_full_size(gpu_available, rmat22):
    """BFS tree at R-MAT scale 22: parents equal the oracle's smallest-parent
    rule and pass the reference's validation rules."""
    c = rmat22
    src = workloads.max_degree_vertex(c)
    with gg.Graph(c) as g:
        d, p, st = g.bfs_tree(src)
    assert oracle.bfs_validate(c.row_offsets, c.col_indices, src, d, p) == 0
    np.testing.assert_array_equal(p, oracle.bfs_parents(c.row_offsets, c.col_indices, d, src))


def test_bfs_tree(gpu_available, kat_graphs):
    """Parents (smallest in-neighbour one level up) equal the oracle's on the
    KATs and on directed / undirected R-MAT, through both BFS paths and with
    device outputs; every tree passes the reference's validation rules."""
    import torch

    for k in kat_graphs:
        with gg.Graph(csr_of(k["row_offsets"], k["col_indices"]), directed=k["directed"]) as g:
            for s in k["bfs"]:
                for push in (False, True):
                    d, p, _ = g.bfs_tree(int(s), push_only=push)
                    assert d.tolist() == k["bfs"][s]
                    np.testing.assert_array_equal(p, oracle.bfs_parents(k["row_offsets"], k["col_indices"], d, int(s)))
    for undirected in (True, False):
        c = gg.gen_rmat(14, 16, undirected=undirected, seed=74)
        srcs = (workloads.max_degree_vertex(c), 5)
        with gg.Graph(c, directed=not undirected) as g:
            for src in srcs:
                d, p, _ = g.bfs_tree(src)
                assert oracle.bfs_validate(c.row_offsets, c.col_indices, src, d, p) == 0
                np.testing.assert_array_equal(p, oracle.bfs_parents(c.row_offsets, c.col_indices, d, src))
                dd = torch.empty(c.n, dtype=torch.int32, device="cuda")
                pd = torch.empty(c.n, dtype=torch.int32, device="cuda")
                g.bfs_tree(src, push_only=True, flags=_abi.GG_OUT_DEVICE, depth_out=dd, pred_out=pd)
                g.synchronize()
                np.testing.assert_array_equal(dd.cpu().numpy(), d)
                np.testing.assert_array_equal(pd.cpu().numpy(), p)


def symmetrized(c):
    """The directed graph read as undirected: every edge in both directions
    (a multigraph where both u->v and v->u exist)."""
    rows = np.repeat(np.arange(c.n), c.degrees())
    su = np.concatenate([rows, c.col_indices])
    sv = np.concatenate([c.col_indices, rows])
    order = np.lexsort((sv, su))
    su, sv = su[order], sv[order]
    ro = np.concatenate([[0], np.cumsum(np.bincount(su, minlength=c.n))]).astype(np.int32)
    return ro, sv.astype(np.int32)


def test_in_and_all_edge_programs(gpu_available):
    """Program variants through the generic engine's in-edge and all-edge
    paths on a directed graph: reverse BFS (EXPAND_IN_EDGES) equals BFS of the
    transpose; weakly connected components (EXPAND_ALL_EDGES) equal CC of the
    symmetrized graph; the reference PageRank over all edges
    (GATHER_ALL_EDGES, degree out + in) matches CPUPR of the symmetrized
    graph within its own tolerance.  Also on a renumbered graph."""
    c = gg.gen_rmat(13, 8, undirected=False, seed=91)
    co, ri = oracle.csc(c.row_offsets, c.col_indices)[:2]
    sro, scol = symmetrized(c)
    indeg = np.diff(co)
    srcs = (0, int(np.argmax(indeg)))
    exp_cc = oracle.cc(sro, scol)[0]
    for reorder in (False, True):
        with gg.Graph(c, directed=True, reorder=reorder) as g:
            for src in srcs:
                d, _ = g.bfs(src, flags=_abi.GG_BFS_REVERSE)
                np.testing.assert_array_equal(d, oracle.bfs(co, ri, src), err_msg=f"reorder={reorder} src={src}")
            lab, _ = g.cc(flags=_abi.GG_CC_WEAK)
            np.testing.assert_array_equal(lab, exp_cc)
            r, st = g.pagerank_reference(flags=_abi.GG_PR_ALL_EDGES)
            exp, _ = oracle.pr_reference(sro, scol)
            assert np.linalg.norm(r.astype(np.float64) - exp) / np.linalg.norm(exp) < 0.02
            with pytest.raises(gg.GGError):
                g.bfs_tree(0, flags=_abi.GG_BFS_REVERSE)
    u = gg.gen_rmat(10, 8, undirected=True, seed=92)
    with gg.Graph(u) as g:
        with pytest.raises(gg.GGError):
            g.cc(flags=_abi.GG_CC_WEAK)
        with pytest.raises(gg.GGError):
            g.pagerank_reference(flags=_abi.GG_PR_ALL_EDGES)


@pytest.mark.parametrize("threshold", [-1, 0, 1 << 30])
def test_gas_path_device_and_host_loops(gpu_available, threshold):
    """The generic engine's device-driven loop (captured iterations, the
    LBS / tiered choice made on the device) and its host-driven loop
    (GG_INSTRUMENT) give the oracle's results for the reference schedules
    (push-only BFS, GAS SSSP, GAS CC), the same BFS statistics (fixed by the
    graph), and stop together under an iteration cap."""
    c = gg.gen_rmat(14, 16, undirected=True, seed=28).with_weights("i32", 1, 64, symmetric=True, seed=29)
    src = workloads.max_degree_vertex(c)
    exp_b = oracle.bfs(c.row_offsets, c.col_indices, src)
    exp_s = oracle.sssp(c.row_offsets, c.col_indices, c.values, src)
    exp_c = oracle.cc(c.row_offsets, c.col_indices)[0]
    with gg.Graph(c) as g:
        g.set_options(threshold=threshold)
        stats = []
        for fl in (0, _abi.GG_INSTRUMENT, 0):  # the second device run replays the cached capture
            d, st = g.bfs(src, push_only=True, flags=fl)
            np.testing.assert_array_equal(d, exp_b)
            stats.append((st["iterations"], st["edges_inspected"], st["reached_edges"]))
            dist, _ = g.sssp(src, flags=fl | _abi.GG_SSSP_GAS)
            np.testing.assert_array_equal(dist, exp_s)
            lab, _ = g.cc(gas=True, flags=fl)
            np.testing.assert_array_equal(lab, exp_c)
        assert stats[0] == stats[1] == stats[2]
        g.set_options(iter_num=2)
        res = [g.bfs(src, push_only=True, flags=fl) for fl in (0, _abi.GG_INSTRUMENT)]
        np.testing.assert_array_equal(res[0][0], res[1][0])
        assert res[0][1]["iterations"] == res[1][1]["iterations"] == 2
        assert res[0][0].max() == 2


def test_gas_long_row_goes_to_lbs(gpu_available):
    """Under the per-algorithm default threshold, a frontier within the
    threshold that holds a row longer than 4 * block_degree (here the hub
    source alone) is expanded by LBS instead of one tiered workgroup; an
    explicit threshold keeps the reference's rule (frontier size alone).
    Results equal the oracle's either way."""
    c = gg.gen_rmat(14, 16, undirected=True, seed=28).with_weights("i32", 1, 64, symmetric=True, seed=29)
    src = workloads.max_degree_vertex(c)
    assert c.degrees()[src] > 4 * 256
    exp_s = oracle.sssp(c.row_offsets, c.col_indices, c.values, src)
    launches = {}
    with gg.Graph(c) as g:
        for th in (-1, 10000):
            g.set_options(threshold=th, wave_degree=64, block_degree=256)
            for fl in (0, _abi.GG_INSTRUMENT):
                g.profile_reset()
                dist, _ = g.sssp(src, flags=fl | _abi.GG_SSSP_GAS)
                np.testing.assert_array_equal(dist, exp_s)
            prof = g.profile()
            launches[th] = {k: v["launches"] for k, v in prof.items()}
    # Same iterations either way; at least the hub's iteration moves from the
    # tiered kernel to the LBS one.
    lbs = {th: launches[th].get("sssp.lbs_expand", 0) for th in launches}
    tier = {th: launches[th].get("sssp.tiered_expand", 0) for th in launches}
    assert lbs[-1] + tier[-1] == lbs[10000] + tier[10000]
    assert lbs[-1] > lbs[10000] and tier[-1] < tier[10000]


@pytest.mark.parametrize("delta", [1, 3, 16])
def test_sssp_pull_skips_settled_rows(gpu_available, delta):
    """A pull skips the gathers of rows whose vertex is at or below the
    pulling bucket's lower edge (no dirty vertex can improve them, weights
    being >= 0).  Zero-weight edges included; bit-exact against the oracle
    in the device- and host-driven loops, and the pull's byte count within
    its model (tiles of settled rows are not visited, gathers of settled
    rows not run)."""
    c = gg.gen_rmat(14, 16, undirected=True, seed=31)
    rows = np.repeat(np.arange(c.n), c.degrees())
    lo, hi = np.minimum(rows, c.col_indices), np.maximum(rows, c.col_indices)
    w = ((lo * 5 + hi * 3) % 9).astype(np.int32)  # symmetric, 0..8, zeros included
    src = workloads.max_degree_vertex(c)
    exp = oracle.sssp(c.row_offsets, c.col_indices, w, src)
    n, m = c.n, c.m
    with gg.Graph(gg.CSR(c.row_offsets, c.col_indices, w), reorder=True) as g:
        for pull_alpha in (3, 1 << 20):
            g.set_options(sssp_delta=delta, pull_alpha=pull_alpha)
            for fl in (0, _abi.GG_INSTRUMENT):
                g.profile_reset()
                dist, st = g.sssp(src, flags=fl)
                np.testing.assert_array_equal(dist, exp)
            pulls = [b for name, _, b in g.profile_launches() if name == "sssp.pull"]
            assert st["pull_iterations"] == len(pulls)
            # 8 bytes per in-edge and per row of the tiles visited (a row split
            # over tiles counts in each), 4 per gather that ran
            assert all(0 < b <= 12.0 * m + 8.0 * (n + (n + m) / 512 + 1) for b in pulls)
            if pull_alpha == 1 << 20:  # pulls whenever the frontier has edges
                assert len(pulls) >= 1


@pytest.mark.parametrize("directed", [False, True])
def test_sssp_pull_rows_odd_n_and_wide_weights(gpu_available, directed):
    """The pull's row prefix and tile skip on a vertex count that is not a
    multiple of 16 (the prefix pass's scalar tail) and with weights too wide
    to pack with the source id (the two-stream pull), pulls forced early and
    late: bit-exact against the oracle."""
    c = gg.gen_erdos_renyi(10007, 80000, seed=41)
    if directed:  # keep an asymmetric subset of the entries
        rows = np.repeat(np.arange(c.n), c.degrees())
        keep = (rows < c.col_indices) | ((rows * 7 + c.col_indices) % 3 == 0)
        ro = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=c.n))]).astype(np.int32)
        c = gg.CSR(ro, c.col_indices[keep].astype(np.int32))
    c = c.with_weights("i32", 0, 1 << 20, symmetric=not directed, seed=42)
    assert c.n % 16 != 0
    src = workloads.max_degree_vertex(c)
    exp = oracle.sssp(c.row_offsets, c.col_indices, c.values, src)
    with gg.Graph(c, directed=directed) as g:
        for delta, pull_alpha in ((1 << 14, 1 << 20), (1 << 16, 3), (1 << 18, 2)):
            g.set_options(sssp_delta=delta, pull_alpha=pull_alpha)
            for fl in (0, _abi.GG_INSTRUMENT):
                dist, st = g.sssp(src, flags=fl)
                np.testing.assert_array_equal(dist, exp)


@pytest.mark.parametrize("undirected", [True, False])
def test_scheduled_replays_follow_and_leave_the_last_schedule(gpu_available, undirected):
    """A device-driven BFS / SSSP first replays the schedule (the kinds of the
    levels / iterations) of the previous run on the same graph, each slot
    holding only its kind's kernels, and goes on with the general loop from
    the first level or iteration whose kind differs (the replay stalls there).
    Sources with different schedules in turn -- the hub, a vertex of a short
    path component, an isolated vertex (one level), a low-degree vertex, the
    hub twice -- must all match the oracle, and BFS statistics must equal the
    host-driven loop's, under options that change which kinds occur."""
    r = gg.gen_rmat(14, 16, undirected=undirected, seed=61)
    L, iso = 12, 3
    n = r.n + L + iso
    rows = [np.repeat(np.arange(r.n), np.diff(r.row_offsets))]
    cols = [r.col_indices.astype(np.int64)]
    a = np.arange(r.n, r.n + L - 1)
    rows += [a, a + 1]
    cols += [a + 1, a]
    row, col = np.concatenate(rows), np.concatenate(cols)
    order = np.lexsort((col, row))
    row, col = row[order], col[order]
    ro = np.zeros(n + 1, np.int64)
    np.add.at(ro, row + 1, 1)
    c = csr_of(np.cumsum(ro), col).with_weights("i32", 1, 64, symmetric=undirected, seed=62)
    deg = np.diff(c.row_offsets)
    hub = workloads.max_degree_vertex(c)
    nz = np.flatnonzero(deg[: r.n] > 0)
    low = int(nz[np.argmin(deg[nz])])
    sources = (hub, hub, r.n + 3, r.n + L, low, hub, hub, r.n + L + iso - 1, low)
    with gg.Graph(c, directed=not undirected) as g:
        for opts in ({}, {"pull_alpha": 2, "sssp_delta": 8}, {"bfs_alpha": 1, "bfs_beta": 1000000, "pull_alpha": 0},
                     {"pull_alpha": 1, "sssp_delta": 64}):
            g.set_options(**gg.default_options().as_dict())
            g.set_options(**opts)
            for src in sources:
                src = int(src)
                exp_b = oracle.bfs(c.row_offsets, c.col_indices, src)
                exp_s = oracle.sssp(c.row_offsets, c.col_indices, c.values, src)
                d, st = g.bfs(src)
                np.testing.assert_array_equal(d, exp_b, err_msg=str(("bfs", opts, src)))
                d2, st2 = g.bfs(src, flags=_abi.GG_INSTRUMENT)
                np.testing.assert_array_equal(d2, exp_b)
                for k in ("iterations", "pull_iterations", "reached_vertices", "reached_edges", "edges_inspected"):
                    assert st[k] == st2[k], (opts, src, k, st[k], st2[k])
                ds, sts = g.sssp(src)
                np.testing.assert_array_equal(ds, exp_s, err_msg=str(("sssp", opts, src)))
                assert sts["reached_vertices"] == int(np.sum(exp_s != INF))
