"""ctypes binding of libgasgraph.so (declarations in include/gasgraph.h).

The shared library is built in-tree (gpugraph_amd/csrc/libgasgraph.so) by
``__graft_entry__.build()`` / ``make -C gpugraph_amd/csrc``.  Loading fails
loudly when it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "csrc", "libgasgraph.so")

GG_OK = 0
GG_ERR_INVALID = 1
GG_ERR_OOM = 2
GG_ERR_OVERFLOW = 3
GG_ERR_HIP = 4
GG_ERR_NOT_CONVERGED = 5
GG_ERR_NO_DEVICE = 6
GG_ERR_IO = 7
GG_ERR_UNSUPPORTED = 8
GG_ERR_INTERNAL = 9

GG_VALUES_NONE = 0
GG_VALUES_I32 = 1
GG_VALUES_F32 = 2

GG_GRAPH_DIRECTED = 1
GG_GRAPH_DEVICE_PTRS = 2
GG_GRAPH_REORDER = 4

GG_OUT_DEVICE = 1
GG_INSTRUMENT = 2
GG_BFS_PUSH_ONLY = 4
GG_CC_HOOK = 8
GG_SSSP_GAS = 16
GG_CC_GAS = 32
GG_BFS_REVERSE = 64
GG_CC_WEAK = 128
GG_PR_ALL_EDGES = 256

GG_TUNE_BFS = 1
GG_TUNE_SSSP = 2
GG_TUNE_CC = 4
GG_TUNE_GAS = 8
GG_TUNE_PAGERANK = 16


class Options(C.Structure):
    _fields_ = [
        ("threshold", C.c_int32),
        ("wave_degree", C.c_int32),
        ("block_degree", C.c_int32),
        ("iter_num", C.c_int32),
        ("bfs_alpha", C.c_int32),
        ("bfs_beta", C.c_int32),
        ("lbs_tile", C.c_int32),
        ("sssp_delta", C.c_int32),
        ("pull_alpha", C.c_int32),
        ("expand_grid", C.c_int32),
        ("xcd_slices", C.c_int32),
        ("gas_expand_grid", C.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class LaunchStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("ms", C.c_double), ("alg_bytes", C.c_double)]


class TuneTrial(C.Structure):
    _fields_ = [
        ("programs", C.c_uint32),
        ("option", C.c_int32),
        ("value", C.c_int32),
        ("ms", C.c_float),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("pull_iterations", C.c_int32),
        ("reached_vertices", C.c_int64),
        ("reached_edges", C.c_int64),
        ("edges_inspected", C.c_int64),
        ("alg_bytes", C.c_double),
        ("gpu_ms", C.c_double),
        ("residual", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchSpec(C.Structure):
    _fields_ = [
        ("num_sssp", C.c_int32),
        ("sssp_sources", C.POINTER(C.c_int32)),
        ("sssp_dist", C.POINTER(C.c_int32)),
        ("num_ppr", C.c_int32),
        ("ppr_seeds", C.POINTER(C.c_int32)),
        ("ppr_rank", C.POINTER(C.c_float)),
        ("damping", C.c_float),
        ("tol_l1", C.c_double),
        ("max_iter", C.c_int32),
        ("num_bfs", C.c_int32),
        ("bfs_sources", C.POINTER(C.c_int32)),
        ("bfs_depth", C.POINTER(C.c_int32)),
    ]


class KernelStat(C.Structure):
    _fields_ = [
        ("name", C.c_char * 48),
        ("launches", C.c_int64),
        ("total_ms", C.c_double),
        ("alg_bytes", C.c_double),
    ]


class Csr(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("m", C.c_int64),
        ("row_offsets", C.POINTER(C.c_int32)),
        ("col_indices", C.POINTER(C.c_int32)),
        ("values", C.c_void_p),
        ("value_dtype", C.c_int32),
    ]


_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_F32P = C.POINTER(C.c_float)

# name -> (restype, argtypes); every symbol declared in include/gasgraph.h.
PROTOTYPES = {
    "gg_last_error": (C.c_char_p, []),
    "gg_version": (C.c_int, []),
    "gg_device_count": (C.c_int, [_I32P]),
    "gg_default_options": (C.c_int, [C.POINTER(Options)]),
    "gg_graph_create": (C.c_int, [C.c_int32, C.c_int64, _P, _P, _P, C.c_int32, C.c_uint32, C.POINTER(_P)]),
    "gg_graph_destroy": (C.c_int, [_P]),
    "gg_graph_set_stream": (C.c_int, [_P, _P]),
    "gg_graph_set_options": (C.c_int, [_P, C.POINTER(Options)]),
    "gg_graph_get_options": (C.c_int, [_P, C.POINTER(Options)]),
    "gg_graph_update_values": (C.c_int, [_P, _P, C.c_uint32]),
    "gg_graph_info": (C.c_int, [_P, _I32P, C.POINTER(C.c_int64), _I32P, _I32P]),
    "gg_graph_device_csr": (C.c_int, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "gg_graph_device_csc": (C.c_int, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "gg_graph_vertex_order": (C.c_int, [_P, C.POINTER(_P), C.POINTER(_P)]),
    "gg_graph_synchronize": (C.c_int, [_P]),
    "gg_bfs": (C.c_int, [_P, C.c_int32, _P, C.c_uint32, C.POINTER(Stats)]),
    "gg_bfs_tree": (C.c_int, [_P, C.c_int32, _P, _P, C.c_uint32, C.POINTER(Stats)]),
    "gg_sssp": (C.c_int, [_P, C.c_int32, _P, C.c_uint32, C.POINTER(Stats)]),
    "gg_cc": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(Stats)]),
    "gg_pagerank_reference": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(Stats)]),
    "gg_pagerank": (C.c_int, [_P, C.c_float, _P, C.c_double, C.c_int32, _P, C.c_uint32, C.POINTER(Stats)]),
    "gg_pagerank_adjoint": (C.c_int, [_P, C.c_float, _P, _P, C.c_double, C.c_int32, _P, _P, C.c_uint32,
                                      C.POINTER(Stats)]),
    "gg_batch_run": (C.c_int, [_P, C.POINTER(BatchSpec), C.c_uint32, C.POINTER(Stats)]),
    "gg_autotune": (C.c_int, [_P, C.c_uint32, C.c_int32, C.c_int32, C.POINTER(Options), C.POINTER(TuneTrial),
                              C.c_int32, _I32P]),
    "gg_profile_reset": (C.c_int, [_P]),
    "gg_profile_get": (C.c_int, [_P, C.POINTER(KernelStat), C.c_int32, _I32P]),
    "gg_profile_launches": (C.c_int, [_P, C.POINTER(LaunchStat), C.c_int32, _I32P]),
    "gg_bfs_noalloc": (C.c_int, [C.c_int32, C.c_int64, _P, _P, C.c_int32, _P, _P, C.POINTER(Stats)]),
    "gg_csr_free": (None, [C.POINTER(Csr)]),
    "gg_read_mtx": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.POINTER(Csr)]),
    "gg_gen_erdos_renyi": (C.c_int, [C.c_int32, C.c_int64, C.c_uint64, C.POINTER(Csr)]),
    "gg_gen_rmat": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_uint64,
                              C.POINTER(Csr)]),
    "gg_gen_lattice": (C.c_int, [C.c_int32, C.c_double, C.c_double, C.c_uint64, C.POINTER(Csr)]),
    "gg_csr_weights": (C.c_int, [C.POINTER(Csr), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64]),
}

_lib = None


class GGError(RuntimeError):
    """A libgasgraph call returned a non-zero status."""

    def __init__(self, code: int, func: str, message: str):
        super().__init__(f"{func} failed with status {code}: {message}")
        self.code = code


def _share_torch_hip_runtime():
    """Make libgasgraph.so bind to the HIP runtime torch ships, if torch is
    installed.  torch's wheel carries its own libamdhip64.so with the same
    soname (libamdhip64.so.7) as /opt/rocm's; loading it first (RTLD_GLOBAL)
    lets the dynamic linker resolve our NEEDED entry to that copy, so one HIP
    runtime serves both and device pointers pass freely between them.  Loaded
    the other way round, the process would hold two runtimes and torch's
    device initialisation fails."""
    import importlib.util

    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return
    for loc in spec.submodule_search_locations:
        hip = os.path.join(loc, "lib", "libamdhip64.so")
        if os.path.exists(hip):
            C.CDLL(hip, mode=C.RTLD_GLOBAL)
            return


def load_library(path: str | None = None):
    """Load libgasgraph.so and attach prototypes; raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"{p} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C gpugraph_amd/csrc`")
    _share_torch_hip_runtime()
    lib = C.CDLL(p)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(code: int, func: str):
    if code != GG_OK:
        msg = load_library().gg_last_error().decode(errors="replace")
        raise GGError(code, func, msg)
