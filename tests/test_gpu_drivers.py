"""The reference's driver sequence compiled against the header-only plugin API
(gpugraph_amd/csrc/drivers/gas_driver.hip, built per program by the
library's Makefile): Config from -p / -c, BuildMarketGraph into a CsrGraph,
CsrProblem(cfg) + FromHostProblem (12-argument form), EnactorVertexCentric
(cfg, verbose), EnactIterativeSearch per source, ExtractResults /
ExtractPreds.  Each driver reads a Matrix Market file written here and its
output file is compared with the oracle: BFS depths, SSSP distances and CC
labels bit-exact, BFS predecessors (MARK_PREDECESSORS) through the
reference's validation rules, PageRank against CPUPR as the reference's
driver checks it (relative L2, PageRank.cu:307-313)."""
import os
import subprocess

import numpy as np
import pytest

import gpugraph_amd as gg
import oracle
from gpugraph_amd import workloads

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVERS = os.path.join(ROOT, "gpugraph_amd", "csrc", "drivers")


def write_mtx(path, ro, col, w=None):
    n = len(ro) - 1
    rows = np.repeat(np.arange(n), np.diff(ro))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate Integer general\n")
        f.write(f"{n} {n} {len(col)}\n")
        vals = np.ones(len(col), np.int64) if w is None else np.asarray(w)
        for u, v, x in zip(rows, col, vals):
            f.write(f"{u + 1} {v + 1} {int(x)}\n")


def run_driver(prog, args, cwd):
    exe = os.path.join(DRIVERS, f"gas_{prog}")
    assert os.path.exists(exe), f"{exe} is not built (make -C gpugraph_amd/csrc)"
    r = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def load(path, dtype):
    return np.loadtxt(path, dtype=dtype, ndmin=1)


def test_drivers_on_kats(gpu_available, kat_graphs, tmp_path):
    for k in kat_graphs:
        d = tmp_path / k["name"]
        d.mkdir()
        ro, col, w = k["row_offsets"], k["col_indices"], k["weights"]
        write_mtx(d / "g.mtx", ro, col, w)
        directed = 1 if k["directed"] else 0
        for s, exp in list(k["bfs"].items())[:2]:
            run_driver("bfs", ["-g", "g.mtx", "-o", "bfs.txt", "-preds", "pred.txt",
                               "-p", f"src={int(s) + 1} directed={directed}"], d)
            depth = load(d / "bfs.txt", np.int32)
            assert depth.tolist() == exp, (k["name"], s)
            pred = load(d / "pred.txt", np.int32)
            assert oracle.bfs_validate(ro, col, int(s), depth, pred) == 0, (k["name"], s)
        for s, exp in list(k["sssp"].items())[:2]:
            run_driver("sssp", ["-g", "g.mtx", "-o", "sssp.txt", "-p", f"src={int(s)} origin=0 directed={directed}"], d)
            assert load(d / "sssp.txt", np.int32).tolist() == exp, (k["name"], s)
        write_mtx(d / "sym.mtx", k["sym_row_offsets"], k["sym_col_indices"])
        run_driver("cc", ["-g", "sym.mtx", "-o", "cc.txt"], d)
        assert load(d / "cc.txt", np.int32).tolist() == k["cc_sym"].tolist(), k["name"]


@pytest.mark.parametrize("threshold", [0, 1 << 30, 300])
def test_drivers_on_rmat(gpu_available, tmp_path, threshold):
    """R-MAT scale 12, directed with weights (BFS, SSSP, PageRank) and
    undirected (CC); both expansion strategies and the per-iteration switch
    between them (threshold)."""
    c = gg.gen_rmat(12, 16, undirected=False, seed=71).with_weights("i32", 1, 64, symmetric=False, seed=72)
    u = gg.gen_rmat(12, 8, undirected=True, seed=73)
    write_mtx(tmp_path / "d.mtx", c.row_offsets, c.col_indices, c.values)
    write_mtx(tmp_path / "u.mtx", u.row_offsets, u.col_indices)
    src = workloads.max_degree_vertex(c)
    (tmp_path / "cfg.txt").write_text(f"src={src}\norigin=0\nthreshold={threshold}\n")
    run_driver("bfs", ["-g", "d.mtx", "-o", "bfs.txt", "-preds", "pred.txt", "-c", "cfg.txt"], tmp_path)
    depth = load(tmp_path / "bfs.txt", np.int32)
    np.testing.assert_array_equal(depth, oracle.bfs(c.row_offsets, c.col_indices, src))
    pred = load(tmp_path / "pred.txt", np.int32)
    assert oracle.bfs_validate(c.row_offsets, c.col_indices, src, depth, pred) == 0
    # Two sources from a file, run one after another; the output is the last one's.
    (tmp_path / "srcs.txt").write_text(f"{src + 1}\n2\n")
    run_driver("sssp", ["-g", "d.mtx", "-o", "sssp.txt", "-sources", "srcs.txt", "-p", f"threshold={threshold}"],
               tmp_path)
    np.testing.assert_array_equal(load(tmp_path / "sssp.txt", np.int32),
                                  oracle.sssp(c.row_offsets, c.col_indices, c.values, 1))
    run_driver("cc", ["-g", "u.mtx", "-o", "cc.txt", "-p", f"threshold={threshold}"], tmp_path)
    np.testing.assert_array_equal(load(tmp_path / "cc.txt", np.int32), oracle.cc(u.row_offsets, u.col_indices)[0])
    out = run_driver("pagerank", ["-g", "d.mtx", "-o", "pr.txt", "-p", f"threshold={threshold}"], tmp_path)
    pr = load(tmp_path / "pr.txt", np.float64)
    exp, _ = oracle.pr_reference(c.row_offsets, c.col_indices)
    assert np.linalg.norm(pr - exp) / np.linalg.norm(exp) < 0.02, out


def test_driver_errors(gpu_available, tmp_path):
    exe = os.path.join(DRIVERS, "gas_sssp")
    r = subprocess.run([exe, "-g", "missing.mtx"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    r = subprocess.run([exe, "-g", "x.mtx", "-p", "no_such_parameter=1"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0 and "not registered" in r.stdout
