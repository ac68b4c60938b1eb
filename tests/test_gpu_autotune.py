"""Schedule options never change results: every compiled LBS tile shape and
a range of expand grids give the oracle's BFS / SSSP / CC bit for bit, and
gg_autotune installs valid options, logs its trials, and leaves results
unchanged."""
import numpy as np
import pytest

import gpugraph_amd as gg
import oracle
from gpugraph_amd import _abi, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat14():
    c = workloads.config2_rmat(scale=14)
    src = workloads.max_degree_vertex(c)
    exp = {
        "bfs": oracle.bfs(c.row_offsets, c.col_indices, src),
        "sssp": oracle.sssp(c.row_offsets, c.col_indices, c.values, src),
        "cc": oracle.cc(c.row_offsets, c.col_indices)[0],
    }
    return c, src, exp


def check_all(g, src, exp, tag):
    for push in (False, True):
        d, _ = g.bfs(src, push_only=push)
        np.testing.assert_array_equal(d, exp["bfs"], err_msg=f"bfs push={push} {tag}")
    for flags in (0, _abi.GG_SSSP_GAS):
        d, _ = g.sssp(src, flags=flags)
        np.testing.assert_array_equal(d, exp["sssp"], err_msg=f"sssp flags={flags} {tag}")
    lab, _ = g.cc()
    np.testing.assert_array_equal(lab, exp["cc"], err_msg=f"cc {tag}")


def test_tile_shapes_and_grids(gpu_available, rmat14):
    c, src, exp = rmat14
    with gg.Graph(c) as g:
        for tile in (512, 1024, 2048):
            for grid in (64, 2048, 8192):
                g.set_options(lbs_tile=tile, expand_grid=grid, threshold=0)
                check_all(g, src, exp, f"tile={tile} grid={grid}")


def test_option_validation(gpu_available, rmat14):
    c, _, _ = rmat14
    with gg.Graph(c) as g:
        for bad in ({"lbs_tile": 768}, {"lbs_tile": 4096}, {"expand_grid": 16}, {"expand_grid": 1 << 17}):
            with pytest.raises(gg.GGError):
                g.set_options(**bad)
        assert g.options().lbs_tile == 1024  # unchanged after rejected calls


def test_autotune(gpu_available, rmat14):
    c, src, exp = rmat14
    with gg.Graph(c) as g:
        best, trials = g.autotune(src=src, reps=2)
        assert trials[0][1] == "baseline" and len(trials) > 20
        assert g.options().as_dict() == best
        assert best["lbs_tile"] in (512, 1024, 2048) and best["wave_degree"] <= best["block_degree"]
        assert all(t[3] > 0 for t in trials)
        check_all(g, src, exp, "tuned")
        with pytest.raises(gg.GGError):
            g.autotune(programs=0, src=src)
        with pytest.raises(gg.GGError):
            g.autotune(src=c.n, reps=1)
    # directed, f32 values: SSSP and CC are dropped, BFS / GAS remain
    cd = workloads.config5_rmat(scale=12)
    with gg.Graph(cd, directed=True) as g:
        best, trials = g.autotune(src=0, reps=1)
        assert trials[0][0] == _abi.GG_TUNE_BFS | _abi.GG_TUNE_GAS | _abi.GG_TUNE_PAGERANK
        assert any(t[1] == "xcd_slices" for t in trials)
        with pytest.raises(gg.GGError):
            g.autotune(programs=_abi.GG_TUNE_CC | _abi.GG_TUNE_SSSP, src=0)
