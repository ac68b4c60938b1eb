"""gg_bfs_noalloc (the allocation-free BFS over a caller's device CSR) against
gg_bfs on the same graphs, through the C ABI: two calls in a row on one GPU,
the second on a smaller graph (the scratch array keeps the first call's
contents), plus the size limits and argument errors."""
import ctypes as C

import numpy as np
import pytest

import gpugraph_amd as gg
import oracle
from gpugraph_amd import _abi

pytestmark = pytest.mark.gpu
MAX_N = 1 << 21
MAX_M = 1 << 26


def noalloc_bfs(lib, torch, csr, src):
    ro = torch.tensor(csr.row_offsets, dtype=torch.int32, device="cuda")
    col = torch.tensor(csr.col_indices, dtype=torch.int32, device="cuda")
    depth = torch.full((csr.n,), 7, dtype=torch.int32, device="cuda")
    st = _abi.Stats()
    torch.cuda.synchronize()
    rc = lib.gg_bfs_noalloc(csr.n, csr.m, C.c_void_p(ro.data_ptr()), C.c_void_p(col.data_ptr()), src,
                            C.c_void_p(depth.data_ptr()), None, C.byref(st))
    _abi.check(rc, "gg_bfs_noalloc")
    return depth.cpu().numpy(), st.as_dict()


def test_noalloc_bfs_matches_gg_bfs(gpu_available):
    import torch

    lib = gg.load_library()
    big = gg.gen_rmat(16, 16, undirected=True, seed=91)
    small = gg.gen_rmat(11, 8, undirected=False, seed=92)
    for csr, directed in ((big, False), (small, True)):
        src = int(np.argmax(csr.degrees()))
        d, st = noalloc_bfs(lib, torch, csr, src)
        with gg.Graph(csr, directed=directed) as g:
            exp, est = g.bfs(src)
        np.testing.assert_array_equal(d, exp)
        np.testing.assert_array_equal(d, oracle.bfs(csr.row_offsets, csr.col_indices, src))
        assert st["reached_vertices"] == est["reached_vertices"]
        assert st["reached_edges"] == est["reached_edges"]


def test_noalloc_bfs_limits_and_errors(gpu_available):
    import torch

    lib = gg.load_library()
    buf = torch.zeros(16, dtype=torch.int32, device="cuda")
    p = C.c_void_p(buf.data_ptr())
    for n, m, src, what in ((MAX_N + 1, 8, 0, "vertices"), (8, MAX_M + 1, 0, "edges"), (8, 8, 8, "out of range")):
        rc = lib.gg_bfs_noalloc(n, m, p, p, src, p, None, None)
        assert rc == _abi.GG_ERR_INVALID
        assert what in lib.gg_last_error().decode()
    assert lib.gg_bfs_noalloc(8, 8, None, p, 0, p, None, None) == _abi.GG_ERR_INVALID
