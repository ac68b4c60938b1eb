"""The C ABI library loads and exports every entry point include/gasgraph.h
declares; argument validation that needs no device.  CPU only (no compute)."""
import ctypes as C
import os
import re

import pytest

import gpugraph_amd as gg
from gpugraph_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gasgraph.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gg_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(_abi.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 28
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(_abi.PROTOTYPES) == declared_symbols()


def test_default_options_and_version():
    o = gg.default_options()
    assert (o.wave_degree, o.block_degree, o.bfs_alpha, o.bfs_beta) == (128, 1024, 14, 24)
    assert (o.lbs_tile, o.sssp_delta, o.pull_alpha, o.expand_grid, o.xcd_slices) == (1024, 0, 0, 2048, -1)
    assert o.gas_expand_grid == 8192
    assert o.threshold == -1 and o.iter_num == 2**31 - 1
    assert gg.load_library().gg_version() == 10000


def test_null_arguments_are_rejected():
    lib = gg.load_library()
    assert lib.gg_device_count(None) == _abi.GG_ERR_INVALID
    assert b"NULL" in lib.gg_last_error()
    assert lib.gg_default_options(None) == _abi.GG_ERR_INVALID
    assert lib.gg_bfs(None, 0, None, 0, None) == _abi.GG_ERR_INVALID
    assert lib.gg_graph_destroy(None) == _abi.GG_OK
    assert lib.gg_autotune(None, _abi.GG_TUNE_BFS, 0, 3, None, None, 0, None) == _abi.GG_ERR_INVALID


def test_graph_create_validation_without_device():
    lib = gg.load_library()
    h = C.c_void_p()
    import numpy as np

    ro = np.array([0, 1, 2], np.int32)
    col = np.array([1, 5], np.int32)
    rc = lib.gg_graph_create(2, 2, ro.ctypes.data, col.ctypes.data, None, 0, 0, C.byref(h))
    cnt = C.c_int32()
    lib.gg_device_count(C.byref(cnt))
    if cnt.value == 0:
        assert rc == _abi.GG_ERR_NO_DEVICE
    else:
        assert rc == _abi.GG_ERR_INVALID  # column 5 out of range
    rc = lib.gg_graph_create(0, 0, ro.ctypes.data, col.ctypes.data, None, 0, 0, C.byref(h))
    assert rc == _abi.GG_ERR_INVALID


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(ImportError):
        _abi.load_library(str(tmp_path / "nope.so"))
