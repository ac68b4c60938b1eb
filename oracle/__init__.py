"""ctypes binding of the CPU oracle (oracle/liboracle.so, built by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product package (gpugraph_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_I = C.POINTER(C.c_int32)
_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_bfs.argtypes = [C.c_int32, _I, _I, C.c_int32, _I]
        L.oracle_bfs_parents.argtypes = [C.c_int32, _I, _I, _I, C.c_int32, _I]
        L.oracle_bfs_validate.argtypes = [C.c_int32, _I, _I, C.c_int32, _I, _I]
        L.oracle_bfs_validate.restype = C.c_int32
        L.oracle_threads.restype = C.c_int32
        L.oracle_bfs_omp.argtypes = [C.c_int32, _I, _I, _I, _I, C.c_int32, _I]
        L.oracle_sssp_omp.argtypes = [C.c_int32, _I, _I, _I, C.c_int32, C.c_int32, _I]
        L.oracle_sssp_selection.argtypes = [C.c_int32, _I, _I, _I, C.c_int32, _I]
        L.oracle_sssp_dijkstra.argtypes = [C.c_int32, _I, _I, _I, C.c_int32, _I]
        L.oracle_cc.argtypes = [C.c_int32, _I, _I, _I]
        L.oracle_cc.restype = C.c_int32
        L.oracle_pr_reference.argtypes = [C.c_int32, _I, _I, _I, _F]
        L.oracle_pr_reference.restype = C.c_int32
        L.oracle_pagerank_f64.argtypes = [C.c_int32, _I, _I, _D, _D, C.c_double, C.c_double, C.c_int32, _D, _D]
        L.oracle_pagerank_f64.restype = C.c_int32
        L.oracle_pagerank_adjoint_f64.argtypes = [C.c_int32, _I, _I, _D, C.c_double, _D, _D, C.c_double, C.c_int32,
                                                  _D, _D, _D]
        L.oracle_pagerank_adjoint_f64.restype = C.c_int32
        L.oracle_cc_omp.argtypes = [C.c_int32, _I, _I, _I]
        L.oracle_cc_omp.restype = C.c_int32
        L.oracle_pagerank_f64_omp.argtypes = [C.c_int32, _I, _I, _I, _D, _D, C.c_double, C.c_double, C.c_int32, _D,
                                              _D]
        L.oracle_pagerank_f64_omp.restype = C.c_int32
        L.oracle_pagerank_fd_f64.argtypes = [C.c_int32, _I, _I, _D, C.c_double, _D, _D, C.c_int32,
                                             C.POINTER(C.c_int64), _D, C.c_double, C.c_int32, _D]
        L.oracle_pagerank_fd_f64.restype = C.c_int32
        _lib = L
    return _lib


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(_I)


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_D)


def bfs(ro, col, src):
    ro, col = np.ascontiguousarray(ro, np.int32), np.ascontiguousarray(col, np.int32)
    out = np.empty(len(ro) - 1, np.int32)
    lib().oracle_bfs(len(ro) - 1, _i(ro), _i(col), int(src), out.ctypes.data_as(_I))
    return out


def bfs_parents(ro, col, depth, src):
    ro, col, depth = (np.ascontiguousarray(x, np.int32) for x in (ro, col, depth))
    out = np.empty(len(ro) - 1, np.int32)
    lib().oracle_bfs_parents(len(ro) - 1, _i(ro), _i(col), _i(depth), int(src), out.ctypes.data_as(_I))
    return out


def bfs_validate(ro, col, src, depth, pred):
    ro, col, depth, pred = (np.ascontiguousarray(x, np.int32) for x in (ro, col, depth, pred))
    return int(lib().oracle_bfs_validate(len(ro) - 1, _i(ro), _i(col), int(src), _i(depth), _i(pred)))


def threads():
    return int(lib().oracle_threads())


def bfs_omp(ro, col, src, in_ro=None, in_col=None):
    """Parallel BFS (OpenMP); in_ro/in_col = in-edges (default: the CSR, for symmetric graphs)."""
    ro, col = np.ascontiguousarray(ro, np.int32), np.ascontiguousarray(col, np.int32)
    iro = ro if in_ro is None else np.ascontiguousarray(in_ro, np.int32)
    icol = col if in_col is None else np.ascontiguousarray(in_col, np.int32)
    out = np.empty(len(ro) - 1, np.int32)
    lib().oracle_bfs_omp(len(ro) - 1, _i(ro), _i(col), _i(iro), _i(icol), int(src), out.ctypes.data_as(_I))
    return out


def sssp_omp(ro, col, w, src, delta=128):
    ro, col, w = (np.ascontiguousarray(x, np.int32) for x in (ro, col, w))
    out = np.empty(len(ro) - 1, np.int32)
    lib().oracle_sssp_omp(len(ro) - 1, _i(ro), _i(col), _i(w), int(src), int(delta), out.ctypes.data_as(_I))
    return out


def sssp(ro, col, w, src, selection=False):
    ro, col, w = (np.ascontiguousarray(x, np.int32) for x in (ro, col, w))
    out = np.empty(len(ro) - 1, np.int32)
    fn = lib().oracle_sssp_selection if selection else lib().oracle_sssp_dijkstra
    fn(len(ro) - 1, _i(ro), _i(col), _i(w), int(src), out.ctypes.data_as(_I))
    return out


def cc(ro, col):
    ro, col = np.ascontiguousarray(ro, np.int32), np.ascontiguousarray(col, np.int32)
    out = np.empty(len(ro) - 1, np.int32)
    sweeps = lib().oracle_cc(len(ro) - 1, _i(ro), _i(col), out.ctypes.data_as(_I))
    return out, sweeps


def cc_omp(ro, col):
    """Min-label CC on all host cores (same labels as cc)."""
    ro, col = np.ascontiguousarray(ro, np.int32), np.ascontiguousarray(col, np.int32)
    out = np.empty(len(ro) - 1, np.int32)
    sweeps = lib().oracle_cc_omp(len(ro) - 1, _i(ro), _i(col), out.ctypes.data_as(_I))
    return out, sweeps


def csc(ro, col):
    """Stable CSC (rows ascending within a column) of a CSR, in numpy."""
    n = len(ro) - 1
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(ro))
    order = np.argsort(col, kind="stable")
    co = np.zeros(n + 1, np.int64)
    np.add.at(co, np.asarray(col, np.int64) + 1, 1)
    return np.cumsum(co).astype(np.int32), rows[order].astype(np.int32), order.astype(np.int32)


def pr_reference(ro, col):
    co, ri, _ = csc(ro, col)
    out = np.empty(len(ro) - 1, np.float32)
    iters = lib().oracle_pr_reference(len(ro) - 1, _i(ro), _i(co), _i(ri), out.ctypes.data_as(_F))
    return out, iters


def pagerank_f64(ro, col, p=None, t=None, d=0.85, tol=1e-12, max_iter=10000):
    n = len(ro) - 1
    r = np.empty(n, np.float64)
    res = C.c_double()
    p = None if p is None else np.ascontiguousarray(p, np.float64)
    t = None if t is None else np.ascontiguousarray(t, np.float64)
    it = lib().oracle_pagerank_f64(n, _i(ro), _i(col), _d(p), _d(t), d, tol, max_iter, r.ctypes.data_as(_D),
                                   C.byref(res))
    return r, it, res.value


def pagerank_f64_omp(ro, co, ri, p_csc=None, t=None, d=0.85, tol=1e-12, max_iter=10000):
    """pagerank_f64 on all host cores over the CSC (co, ri); p_csc = transition
    weights in CSC slot order (None: 1/outdeg)."""
    n = len(ro) - 1
    r = np.empty(n, np.float64)
    res = C.c_double()
    p_csc = None if p_csc is None else np.ascontiguousarray(p_csc, np.float64)
    t = None if t is None else np.ascontiguousarray(t, np.float64)
    it = lib().oracle_pagerank_f64_omp(n, _i(ro), _i(co), _i(ri), _d(p_csc), _d(t), d, tol, max_iter,
                                       r.ctypes.data_as(_D), C.byref(res))
    return r, it, res.value


def pagerank_adjoint_f64(ro, col, w, r, c, d=0.85, tol=1e-14, max_iter=10000):
    n = len(ro) - 1
    m = len(col)
    lam = np.empty(n, np.float64)
    gw = np.zeros(m, np.float64)
    gt = np.empty(n, np.float64)
    w = None if w is None else np.ascontiguousarray(w, np.float64)
    r = np.ascontiguousarray(r, np.float64)
    c = np.ascontiguousarray(c, np.float64)
    it = lib().oracle_pagerank_adjoint_f64(n, _i(ro), _i(col), _d(w), d, _d(r), _d(c), tol, max_iter,
                                           lam.ctypes.data_as(_D), gw.ctypes.data_as(_D), gt.ctypes.data_as(_D))
    return lam, gw, gt, it


def pagerank_fd_f64(ro, col, w, r, c, edges, h, d=0.85, tol_rel=1e-10, max_iter=10000):
    """Central differences of L = c . r(w) on the given CSR edges (see
    oracle_pagerank_fd_f64); r = the FP64 fixed point at w."""
    n = len(ro) - 1
    w = np.ascontiguousarray(w, np.float64)
    r = np.ascontiguousarray(r, np.float64)
    c = np.ascontiguousarray(c, np.float64)
    edges = np.ascontiguousarray(edges, np.int64)
    h = np.ascontiguousarray(h, np.float64)
    fd = np.empty(len(edges), np.float64)
    sweeps = lib().oracle_pagerank_fd_f64(n, _i(ro), _i(col), _d(w), d, _d(r), _d(c), len(edges),
                                          edges.ctypes.data_as(C.POINTER(C.c_int64)), _d(h), tol_rel, max_iter,
                                          fd.ctypes.data_as(_D))
    return fd, sweeps
