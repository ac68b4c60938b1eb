"""gpugraph_amd: MI355X (gfx950) Gather-Apply-Scatter graph engine.

Python host API over libgasgraph.so (C ABI in include/gasgraph.h).  The
compute path is HIP only; importing the package does not need a GPU, but
every run call does and fails loudly without one.

    import gpugraph_amd as gg
    csr = gg.gen_rmat(scale=16, edge_factor=16, undirected=True, seed=2)
    with gg.Graph(csr) as g:
        depth, stats = g.bfs(src=0)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from ._abi import GGError, load_library

__all__ = [
    "CSR", "Graph", "GGError", "load_library", "gen_erdos_renyi", "gen_rmat", "gen_lattice", "read_mtx",
    "SSSP_INF", "default_options",
]

SSSP_INF = 100000000  # Algorithms/SSSP/sssp.h:43 INIT_VALUE


@dataclass
class CSR:
    """Host CSR graph (int32 offsets / column ids, optional int32/f32 values)."""

    row_offsets: np.ndarray
    col_indices: np.ndarray
    values: np.ndarray | None = None
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.row_offsets.shape[0] - 1)

    @property
    def m(self) -> int:
        return int(self.col_indices.shape[0])

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_offsets)

    def with_weights(self, dtype: str, lo: int = 1, hi: int = 64, symmetric: bool = True, seed: int = 0) -> "CSR":
        """Attach seeded edge values (int32 uniform [lo,hi] or f32 (k+1)/2^24)."""
        code = _abi.GG_VALUES_I32 if dtype == "i32" else _abi.GG_VALUES_F32
        lib = load_library()
        ro = np.ascontiguousarray(self.row_offsets, dtype=np.int32)
        col = np.ascontiguousarray(self.col_indices, dtype=np.int32)
        c = _abi.Csr()
        c.n, c.m = self.n, self.m
        c.row_offsets = ro.ctypes.data_as(_abi._I32P)
        c.col_indices = col.ctypes.data_as(_abi._I32P)
        c.values, c.value_dtype = None, 0
        _abi.check(lib.gg_csr_weights(C.byref(c), code, lo, hi, int(symmetric), seed), "gg_csr_weights")
        ct = C.c_int32 if code == _abi.GG_VALUES_I32 else C.c_float
        vals = np.ctypeslib.as_array(C.cast(c.values, C.POINTER(ct)), shape=(max(self.m, 1),))[:self.m].copy()
        # Only the values buffer belongs to the library; detach the numpy-backed arrays before freeing.
        c.row_offsets = C.cast(None, _abi._I32P)
        c.col_indices = C.cast(None, _abi._I32P)
        lib.gg_csr_free(C.byref(c))
        return CSR(self.row_offsets, self.col_indices, vals, dict(self.meta))


def _from_c(c: "_abi.Csr", owned: bool = True) -> CSR:
    n, m = int(c.n), int(c.m)
    ro = np.ctypeslib.as_array(c.row_offsets, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(c.col_indices, shape=(max(m, 1),))[:m].copy() if m else np.zeros(0, np.int32)
    vals = None
    if c.values and c.value_dtype:
        ct = C.c_int32 if c.value_dtype == _abi.GG_VALUES_I32 else C.c_float
        arr = C.cast(c.values, C.POINTER(ct))
        vals = np.ctypeslib.as_array(arr, shape=(max(m, 1),))[:m].copy()
    if owned:
        load_library().gg_csr_free(C.byref(c))
    return CSR(ro.astype(np.int32), col.astype(np.int32), vals)


def gen_erdos_renyi(n: int, num_pairs: int, seed: int) -> CSR:
    c = _abi.Csr()
    _abi.check(load_library().gg_gen_erdos_renyi(n, num_pairs, seed, C.byref(c)), "gg_gen_erdos_renyi")
    out = _from_c(c)
    out.meta.update(kind="erdos_renyi", n=n, pairs=num_pairs, seed=seed, undirected=True)
    return out


def gen_rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             undirected: bool = True, seed: int = 0) -> CSR:
    cs = _abi.Csr()
    _abi.check(load_library().gg_gen_rmat(scale, edge_factor, a, b, c, int(undirected), seed, C.byref(cs)),
               "gg_gen_rmat")
    out = _from_c(cs)
    out.meta.update(kind="rmat", scale=scale, edge_factor=edge_factor, seed=seed, undirected=undirected)
    return out


def gen_lattice(side: int, keep_p: float = 0.6, long_range_frac: float = 0.001, seed: int = 0) -> CSR:
    c = _abi.Csr()
    _abi.check(load_library().gg_gen_lattice(side, keep_p, long_range_frac, seed, C.byref(c)), "gg_gen_lattice")
    out = _from_c(c)
    out.meta.update(kind="lattice", side=side, seed=seed, undirected=True)
    return out


def read_mtx(path: str, undirected: bool = False, values: str | None = "i32") -> CSR:
    code = {None: _abi.GG_VALUES_NONE, "i32": _abi.GG_VALUES_I32, "f32": _abi.GG_VALUES_F32}[values]
    c = _abi.Csr()
    _abi.check(load_library().gg_read_mtx(path.encode(), int(undirected), code, C.byref(c)), "gg_read_mtx")
    out = _from_c(c)
    out.meta.update(kind="mtx", path=path, undirected=undirected)
    return out


def default_options() -> "_abi.Options":
    o = _abi.Options()
    _abi.check(load_library().gg_default_options(C.byref(o)), "gg_default_options")
    return o


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def _dev_ptr(x) -> int:
    """Device address of a torch tensor (must be contiguous) or a raw int."""
    if hasattr(x, "data_ptr"):
        if not x.is_contiguous():
            raise ValueError("device tensors must be contiguous")
        return int(x.data_ptr())
    return int(x)


class Graph:
    """A graph resident on the current HIP device (gg_graph handle)."""

    def __init__(self, csr: CSR, directed: bool = False, reorder: bool = False):
        self.lib = load_library()
        self.csr = csr
        self.n, self.m = csr.n, csr.m
        self.directed = directed
        ro = np.ascontiguousarray(csr.row_offsets, dtype=np.int32)
        col = np.ascontiguousarray(csr.col_indices, dtype=np.int32)
        vals, dt = None, _abi.GG_VALUES_NONE
        if csr.values is not None:
            if csr.values.dtype == np.int32:
                vals, dt = np.ascontiguousarray(csr.values), _abi.GG_VALUES_I32
            else:
                vals, dt = np.ascontiguousarray(csr.values, dtype=np.float32), _abi.GG_VALUES_F32
        self.value_dtype = dt
        # Bumped by every update_values: callers that keep results computed
        # from the values (autograd) check it before reusing the graph.
        self.values_version = 0
        h = C.c_void_p()
        _abi.check(self.lib.gg_graph_create(self.n, self.m, _ptr(ro), _ptr(col), _ptr(vals) if vals is not None else None,
                                            dt, (_abi.GG_GRAPH_DIRECTED if directed else 0)
                                            | (_abi.GG_GRAPH_REORDER if reorder else 0), C.byref(h)),
                   "gg_graph_create")
        self.h = h

    # -- lifecycle -----------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.gg_graph_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- options / profiling -------------------------------------------------
    def options(self) -> "_abi.Options":
        o = _abi.Options()
        _abi.check(self.lib.gg_graph_get_options(self.h, C.byref(o)), "gg_graph_get_options")
        return o

    def set_options(self, **kw):
        o = self.options()
        for k, v in kw.items():
            setattr(o, k, v)
        _abi.check(self.lib.gg_graph_set_options(self.h, C.byref(o)), "gg_graph_set_options")

    def autotune(self, programs: int | None = None, src: int = 0, reps: int = 3, max_trials: int = 256):
        """Search the schedule options on this graph (gg_autotune) and install
        the best.  programs: GG_TUNE_* bits (default: every program this
        graph supports).  Returns (best options dict, trials) where trials is
        a list of (programs, option name or "baseline", value, ms)."""
        if programs is None:
            programs = _abi.GG_TUNE_BFS | _abi.GG_TUNE_GAS | _abi.GG_TUNE_PAGERANK
            if self.value_dtype == _abi.GG_VALUES_I32:
                programs |= _abi.GG_TUNE_SSSP
            if not self.directed:
                programs |= _abi.GG_TUNE_CC
        best = _abi.Options()
        arr = (_abi.TuneTrial * max(max_trials, 1))()
        cnt = C.c_int32()
        _abi.check(self.lib.gg_autotune(self.h, programs, src, reps, C.byref(best), arr, max_trials, C.byref(cnt)),
                   "gg_autotune")
        names = [f for f, _ in _abi.Options._fields_]
        label = {-1: "baseline", -2: "final"}
        trials = [(t.programs, label[t.option] if t.option < 0 else names[t.option], t.value, float(t.ms))
                  for t in arr[:cnt.value]]
        return best.as_dict(), trials

    def set_stream(self, stream_ptr: int | None):
        _abi.check(self.lib.gg_graph_set_stream(self.h, C.c_void_p(stream_ptr) if stream_ptr else None),
                   "gg_graph_set_stream")

    def profile_launches(self) -> list:
        """Per-launch (name, ms, alg_bytes) of instrumented runs since the last reset."""
        cnt = C.c_int32()
        _abi.check(self.lib.gg_profile_launches(self.h, None, 0, C.byref(cnt)), "gg_profile_launches")
        arr = (_abi.LaunchStat * max(cnt.value, 1))()
        _abi.check(self.lib.gg_profile_launches(self.h, arr, cnt.value, C.byref(cnt)), "gg_profile_launches")
        return [(a.name.decode(), a.ms, a.alg_bytes) for a in arr[:cnt.value]]

    def profile(self) -> dict:
        cnt = C.c_int32()
        _abi.check(self.lib.gg_profile_get(self.h, None, 0, C.byref(cnt)), "gg_profile_get")
        arr = (_abi.KernelStat * max(cnt.value, 1))()
        _abi.check(self.lib.gg_profile_get(self.h, arr, cnt.value, C.byref(cnt)), "gg_profile_get")
        return {arr[i].name.decode(): dict(launches=arr[i].launches, total_ms=arr[i].total_ms,
                                           alg_bytes=arr[i].alg_bytes) for i in range(cnt.value)}

    def profile_reset(self):
        _abi.check(self.lib.gg_profile_reset(self.h), "gg_profile_reset")

    def synchronize(self):
        _abi.check(self.lib.gg_graph_synchronize(self.h), "gg_graph_synchronize")

    # -- programs ------------------------------------------------------------
    def bfs(self, src: int, push_only: bool = False, flags: int = 0, out=None):
        st = _abi.Stats()
        f = flags | (_abi.GG_BFS_PUSH_ONLY if push_only else 0)
        out, ptr = self._out(out, np.int32, f)
        _abi.check(self.lib.gg_bfs(self.h, src, ptr, f, C.byref(st)), "gg_bfs")
        return out, st.as_dict()

    def bfs_tree(self, src: int, push_only: bool = False, flags: int = 0, depth_out=None, pred_out=None):
        """BFS depths and parents (gg_bfs_tree): pred[src] = src, pred[v] =
        the smallest in-neighbour one level up, -1 where unreached."""
        st = _abi.Stats()
        f = flags | (_abi.GG_BFS_PUSH_ONLY if push_only else 0)
        depth, dptr = self._out(depth_out, np.int32, f)
        pred, pptr = self._out(pred_out, np.int32, f)
        _abi.check(self.lib.gg_bfs_tree(self.h, src, dptr, pptr, f, C.byref(st)), "gg_bfs_tree")
        return depth, pred, st.as_dict()

    def sssp(self, src: int, flags: int = 0, out=None):
        st = _abi.Stats()
        out, ptr = self._out(out, np.int32, flags)
        _abi.check(self.lib.gg_sssp(self.h, src, ptr, flags, C.byref(st)), "gg_sssp")
        return out, st.as_dict()

    def cc(self, hook: bool = False, gas: bool = False, flags: int = 0, out=None):
        """CC labels (gg_cc): label propagation (default), through the generic
        GAS path (gas=True) or by hooking (hook=True)."""
        st = _abi.Stats()
        f = flags | (_abi.GG_CC_HOOK if hook else 0) | (_abi.GG_CC_GAS if gas else 0)
        out, ptr = self._out(out, np.int32, f)
        _abi.check(self.lib.gg_cc(self.h, ptr, f, C.byref(st)), "gg_cc")
        return out, st.as_dict()

    def pagerank_reference(self, flags: int = 0, out=None):
        st = _abi.Stats()
        out, ptr = self._out(out, np.float32, flags)
        _abi.check(self.lib.gg_pagerank_reference(self.h, ptr, flags, C.byref(st)), "gg_pagerank_reference")
        return out, st.as_dict()

    def pagerank(self, damping: float = 0.85, teleport=None, tol_l1: float = 1e-6, max_iter: int = 1000,
                 flags: int = 0, out=None):
        """FP32 PageRank (gg_pagerank).  With GG_OUT_DEVICE, `teleport` and
        `out` are device pointers (ints) or torch tensors."""
        st = _abi.Stats()
        out, ptr = self._out(out, np.float32, flags)
        tp = None
        if teleport is not None:
            if flags & _abi.GG_OUT_DEVICE:
                tp = C.c_void_p(_dev_ptr(teleport))
            else:
                teleport = np.ascontiguousarray(teleport, dtype=np.float32)
                tp = _ptr(teleport)
        rc = self.lib.gg_pagerank(self.h, damping, tp, tol_l1, max_iter, ptr, flags, C.byref(st))
        _abi.check(rc, "gg_pagerank")
        return out, st.as_dict()

    def pagerank_adjoint(self, rank, c, damping: float = 0.85, tol_l1: float = 1e-7, max_iter: int = 1000,
                         flags: int = 0, grad_w=None, grad_t=None):
        """Adjoint of pagerank for L = c . r at the stored ranks `rank`.
        Host arrays by default; with GG_OUT_DEVICE every array argument is a
        device pointer / torch tensor and grad_w / grad_t must be given."""
        st = _abi.Stats()
        if flags & _abi.GG_OUT_DEVICE:
            args = [C.c_void_p(_dev_ptr(x)) if x is not None else None for x in (rank, c, grad_w, grad_t)]
            _abi.check(self.lib.gg_pagerank_adjoint(self.h, damping, args[0], args[1], tol_l1, max_iter, args[2],
                                                    args[3], flags, C.byref(st)), "gg_pagerank_adjoint")
            return grad_w, grad_t, st.as_dict()
        rank = np.ascontiguousarray(rank, dtype=np.float32)
        c = np.ascontiguousarray(c, dtype=np.float32)
        gw = np.empty(self.m, np.float32)
        gt = np.empty(self.n, np.float32)
        _abi.check(self.lib.gg_pagerank_adjoint(self.h, damping, _ptr(rank), _ptr(c), tol_l1, max_iter, _ptr(gw),
                                                _ptr(gt), flags, C.byref(st)), "gg_pagerank_adjoint")
        return gw, gt, st.as_dict()

    def update_values(self, values, device: bool = False):
        """Replace the edge values (same dtype): host array, or a device
        pointer / torch tensor with device=True."""
        if device:
            _abi.check(self.lib.gg_graph_update_values(self.h, C.c_void_p(_dev_ptr(values)), _abi.GG_OUT_DEVICE),
                       "gg_graph_update_values")
        else:
            dt = np.int32 if self.value_dtype == _abi.GG_VALUES_I32 else np.float32
            v = np.ascontiguousarray(values, dtype=dt)
            _abi.check(self.lib.gg_graph_update_values(self.h, _ptr(v), 0), "gg_graph_update_values")
        self.values_version += 1

    def batch(self, sssp_sources=(), ppr_seeds=(), damping: float = 0.85, tol_l1: float = 1e-6,
              max_iter: int = 1000, flags: int = 0, dist_out=None, rank_out=None, bfs_sources=(), depth_out=None):
        """Batched SSSP / personalised PageRank / BFS (gg_batch_run).  Results
        are instance-major: dist[i][v], rank[j][v], depth[k][v].  With
        GG_OUT_DEVICE in flags, dist_out / rank_out / depth_out are device
        buffers (contiguous tensors or raw addresses) of len(sources) x n int32
        / len(seeds) x n float32 / len(bfs_sources) x n int32, and are returned
        in place of host arrays.  Returns (dist, rank, stats), or (dist, rank,
        depth, stats) when bfs_sources are given."""
        src = np.ascontiguousarray(sssp_sources, dtype=np.int32)
        seeds = np.ascontiguousarray(ppr_seeds, dtype=np.int32)
        bsrc = np.ascontiguousarray(bfs_sources, dtype=np.int32)
        spec = _abi.BatchSpec()
        if flags & _abi.GG_OUT_DEVICE:
            if (len(src) and dist_out is None) or (len(seeds) and rank_out is None) or (
                    len(bsrc) and depth_out is None):
                raise ValueError("GG_OUT_DEVICE needs dist_out / rank_out / depth_out device buffers")
            dist, rank, depth = dist_out, rank_out, depth_out
            spec.sssp_dist = C.cast(C.c_void_p(_dev_ptr(dist_out) if len(src) else 0), _abi._I32P)
            spec.ppr_rank = C.cast(C.c_void_p(_dev_ptr(rank_out) if len(seeds) else 0), _abi._F32P)
            spec.bfs_depth = C.cast(C.c_void_p(_dev_ptr(depth_out) if len(bsrc) else 0), _abi._I32P)
        else:
            dist = np.empty((len(src), self.n), np.int32)
            rank = np.empty((len(seeds), self.n), np.float32)
            depth = np.empty((len(bsrc), self.n), np.int32)
            spec.sssp_dist = dist.ctypes.data_as(_abi._I32P)
            spec.ppr_rank = rank.ctypes.data_as(_abi._F32P)
            spec.bfs_depth = depth.ctypes.data_as(_abi._I32P)
        spec.num_bfs = len(bsrc)
        spec.bfs_sources = bsrc.ctypes.data_as(_abi._I32P)
        spec.num_sssp = len(src)
        spec.sssp_sources = src.ctypes.data_as(_abi._I32P)
        spec.num_ppr = len(seeds)
        spec.ppr_seeds = seeds.ctypes.data_as(_abi._I32P)
        spec.damping = damping
        spec.tol_l1 = tol_l1
        spec.max_iter = max_iter
        st = _abi.Stats()
        _abi.check(self.lib.gg_batch_run(self.h, C.byref(spec), flags, C.byref(st)), "gg_batch_run")
        if len(bsrc):
            return dist, rank, depth, st.as_dict()
        return dist, rank, st.as_dict()

    def _out(self, out, dtype, flags):
        if flags & _abi.GG_OUT_DEVICE:
            if out is None:
                raise ValueError("GG_OUT_DEVICE needs a device pointer (int) or tensor in out")
            return out, C.c_void_p(_dev_ptr(out))
        if out is None:
            out = np.empty(self.n, dtype)
        return out, _ptr(out)
