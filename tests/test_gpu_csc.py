"""The CSC built on the device at graph creation (stable radix sort of
(column, edge id)) equals the stable counting-sort transpose computed here
with numpy, entry for entry: column offsets, source rows ascending within a
column, and the CSC->CSR edge ids.  Device-side CSR validation rejects bad
inputs."""
import ctypes as C

import numpy as np
import pytest

import gpugraph_amd as gg
from gpugraph_amd import _abi

pytestmark = pytest.mark.gpu
D2H = 2  # hipMemcpyDeviceToHost


def fetch(ptr, count):
    out = np.empty(max(count, 1), np.int32)
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    assert hip.hipMemcpy(out.ctypes.data, ptr, 4 * count, D2H) == 0
    return out[:count]


def device_csc(g):
    co, ri, eid = C.c_void_p(), C.c_void_p(), C.c_void_p()
    _abi.check(g.lib.gg_graph_device_csc(g.h, C.byref(co), C.byref(ri), C.byref(eid)), "gg_graph_device_csc")
    return fetch(co.value, g.n + 1), fetch(ri.value, g.m), (fetch(eid.value, g.m) if eid.value else None)


def numpy_csc(c):
    eid = np.argsort(c.col_indices, kind="stable").astype(np.int32)
    rows = np.repeat(np.arange(c.n, dtype=np.int32), np.diff(c.row_offsets))
    co = np.searchsorted(c.col_indices[eid], np.arange(c.n + 1), side="left").astype(np.int32)
    return co, rows[eid], eid


def graphs():
    yield gg.gen_rmat(14, 16, undirected=False, seed=71)
    # duplicates, self-loops, empty rows and columns, n not a power of two
    rng = np.random.default_rng(72)
    n = 1000
    deg = rng.integers(0, 9, n)
    deg[::7] = 0
    ro = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    col = rng.integers(0, n // 2, ro[-1]).astype(np.int32)
    yield gg.CSR(ro, col)
    yield gg.CSR(np.zeros(6, np.int32), np.zeros(0, np.int32))  # no edges


def test_device_csc_matches_stable_transpose(gpu_available):
    for c in graphs():
        with gg.Graph(c, directed=True) as g:
            co, ri, eid = device_csc(g)
        eco, eri, eeid = numpy_csc(c)
        np.testing.assert_array_equal(co, eco)
        np.testing.assert_array_equal(ri, eri)
        np.testing.assert_array_equal(eid, eeid)


def test_undirected_view_is_the_csr(gpu_available):
    c = gg.gen_rmat(10, 8, undirected=True, seed=73)
    with gg.Graph(c) as g:
        co, ri, eid = device_csc(g)
    np.testing.assert_array_equal(co, c.row_offsets)
    np.testing.assert_array_equal(ri, c.col_indices)
    assert eid is None


def test_device_validation_rejects_bad_csr(gpu_available):
    ro = np.array([0, 2, 1, 3], np.int32)  # decreasing offsets
    with pytest.raises(gg.GGError) as e:
        gg.Graph(gg.CSR(ro, np.zeros(3, np.int32)), directed=True)
    assert e.value.code == _abi.GG_ERR_INVALID
    ro = np.array([0, 2, 3, 3], np.int32)
    with pytest.raises(gg.GGError) as e:
        gg.Graph(gg.CSR(ro, np.array([0, 3, 1], np.int32)), directed=True)  # column 3 out of range
    assert b"out of range" in gg.load_library().gg_last_error()
    ro = np.array([1, 2, 3], np.int32)  # does not start at 0
    with pytest.raises(gg.GGError):
        gg.Graph(gg.CSR(ro, np.zeros(3, np.int32)))
