"""Host graph builders: structural invariants, determinism pins (the GPU box
must rebuild exactly the graphs the oracle sees), the MTX reader's reference
semantics, and the workload definitions.  CPU only."""
import hashlib
import os

import numpy as np
import pytest

import gpugraph_amd as gg
from gpugraph_amd import workloads

REF_SMALL = os.path.join(os.path.dirname(__file__), "golden", "kat_graphs.json")


def digest(c):
    x = hashlib.sha256()
    x.update(c.row_offsets.tobytes())
    x.update(c.col_indices.tobytes())
    if c.values is not None:
        x.update(c.values.tobytes())
    return x.hexdigest()[:16]


def check_csr(c, symmetric, simple=True):
    ro, col = c.row_offsets, c.col_indices
    assert ro[0] == 0 and ro[-1] == len(col)
    assert np.all(np.diff(ro) >= 0)
    assert col.min(initial=0) >= 0 and col.max(initial=0) < c.n
    rows = np.repeat(np.arange(c.n), np.diff(ro))
    if simple:
        assert not np.any(rows == col), "self-loop"
        key = rows.astype(np.int64) * c.n + col
        assert np.all(np.diff(key) > 0), "rows must be sorted without duplicates"
    if symmetric:
        fwd = np.sort(rows.astype(np.int64) * c.n + col)
        rev = np.sort(col.astype(np.int64) * c.n + rows)
        np.testing.assert_array_equal(fwd, rev)


def test_erdos_renyi_exact_pairs():
    c = gg.gen_erdos_renyi(4096, 16384, 1)
    assert c.n == 4096 and c.m == 32768
    check_csr(c, symmetric=True)


@pytest.mark.parametrize("undirected", [True, False])
def test_rmat_invariants(undirected):
    c = gg.gen_rmat(12, 16, undirected=undirected, seed=2)
    assert c.n == 4096
    check_csr(c, symmetric=undirected)
    deg = c.degrees()
    assert deg.max() > 20 * deg.mean()  # skewed


def test_lattice_invariants():
    c = gg.gen_lattice(64, 0.6, 0.01, 3)
    check_csr(c, symmetric=True)
    # most edges are lattice bonds: |u - v| in {1, side}
    rows = np.repeat(np.arange(c.n), c.degrees())
    diff = np.abs(rows - c.col_indices)
    assert np.mean((diff == 1) | (diff == 64)) > 0.95


def test_weights_ranges_and_symmetry():
    c = gg.gen_erdos_renyi(512, 2000, 5).with_weights("i32", 1, 64, True, 7)
    assert c.values.dtype == np.int32 and c.values.min() >= 1 and c.values.max() <= 64
    rows = np.repeat(np.arange(c.n), c.degrees())
    wmap = {(int(u), int(v)): int(w) for u, v, w in zip(rows, c.col_indices, c.values)}
    assert all(wmap[(v, u)] == w for (u, v), w in wmap.items())
    f = gg.gen_rmat(10, 8, undirected=False, seed=5).with_weights("f32", symmetric=False, seed=13)
    assert f.values.dtype == np.float32 and f.values.min() > 0 and f.values.max() <= 1.0
    # (k+1)/2^24 is exact in FP32
    k = f.values.astype(np.float64) * 2**24
    np.testing.assert_array_equal(k, np.round(k))


def test_determinism_pins():
    assert digest(gg.gen_erdos_renyi(4096, 16384, 1)) == "b78d3f2fbc1d92b9"
    assert digest(gg.gen_erdos_renyi(4096, 16384, 1).with_weights("i32", 1, 64, True, 11)) == "8c304727166a520a"
    assert digest(gg.gen_rmat(12, 16, undirected=True, seed=2)) == "4d2088cbfcf12f11"
    assert digest(gg.gen_rmat(12, 16, undirected=False, seed=4)) == "85914e13119abcdc"
    assert digest(gg.gen_rmat(12, 16, undirected=False, seed=5).with_weights("f32", symmetric=False,
                                                                             seed=13)) == "8fb4aae67f0415e1"
    assert digest(gg.gen_lattice(64, 0.6, 0.001, 3)) == "7aa68ff8f3959287"


def test_thread_count_independence(monkeypatch):
    a = gg.gen_rmat(14, 16, undirected=True, seed=9)
    monkeypatch.setenv("GG_HOST_THREADS", "1")
    b = gg.gen_rmat(14, 16, undirected=True, seed=9)
    assert digest(a) == digest(b)


def test_mtx_reader_matches_reference_semantics(kat_graphs, tmp_path):
    # small.mtx restated: 1-based ids, (row, col) order, values kept.
    g = next(x for x in kat_graphs if x["name"] == "small")
    lines = ["%%MatrixMarket matrix coordinate Integer general", "7 7 11"]
    rows = np.repeat(np.arange(g["n"]), np.diff(g["row_offsets"]))
    order = np.random.default_rng(0).permutation(len(rows))  # file order must not matter
    for i in order:
        lines.append(f"{rows[i] + 1} {g['col_indices'][i] + 1} {g['weights'][i]}")
    p = tmp_path / "small.mtx"
    p.write_text("\n".join(lines) + "\n")
    c = gg.read_mtx(str(p), undirected=False, values="i32")
    np.testing.assert_array_equal(c.row_offsets, g["row_offsets"])
    np.testing.assert_array_equal(c.col_indices, g["col_indices"])
    np.testing.assert_array_equal(c.values, g["weights"])
    # undirected: every line mirrored, duplicates and self-loops kept
    p2 = tmp_path / "dup.mtx"
    p2.write_text("%%MatrixMarket matrix coordinate Integer general\n3 3 3\n1 2 5\n1 2 5\n3 3 1\n")
    c2 = gg.read_mtx(str(p2), undirected=True, values="i32")
    assert c2.row_offsets.tolist() == [0, 2, 4, 6]
    assert c2.col_indices.tolist() == [1, 1, 0, 0, 2, 2]


def test_mtx_reader_errors(tmp_path):
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate Integer general\n3 4 1\n1 2 1\n")
    with pytest.raises(gg.GGError) as e:
        gg.read_mtx(str(bad))
    assert e.value.code == 7  # GG_ERR_IO: not square
    short = tmp_path / "short.mtx"
    short.write_text("%%MatrixMarket matrix coordinate Integer general\n3 3 2\n1 2 1\n")
    with pytest.raises(gg.GGError):
        gg.read_mtx(str(short))
    with pytest.raises(gg.GGError):
        gg.read_mtx(str(tmp_path / "missing.mtx"))


def test_generator_argument_errors():
    with pytest.raises(gg.GGError):
        gg.gen_rmat(0, 16)
    with pytest.raises(gg.GGError):
        gg.gen_erdos_renyi(4, 100, 1)  # more pairs than exist
    with pytest.raises(gg.GGError):
        gg.gen_lattice(1)


def test_workload_small_configs():
    c1 = workloads.config1_er()
    assert c1.n == 4096 and c1.m == 32768
    src = workloads.max_degree_vertex(c1)
    assert c1.degrees()[src] == c1.degrees().max()
    assert src == int(np.argmax(c1.degrees()))  # ties -> lowest id
