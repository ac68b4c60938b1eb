"""Differentiable PageRank for PyTorch over libgasgraph.so (ctypes).

    g = gg.Graph(csr_with_f32_values, directed=True)
    w = torch.tensor(csr.values, device="cuda", requires_grad=True)
    t = torch.full((n,), 1.0 / n, device="cuda", requires_grad=True)
    r = pagerank(g, w, t)          # forward: gg_pagerank on the device
    (c * r).sum().backward()       # backward: gg_pagerank_adjoint

forward() loads the raw edge weights into the resident graph (the library
normalises each source's out-weights), runs the compensated FP32 forward and
saves the ranks and the weights; backward() hands the saved ranks and the
incoming gradient to the adjoint kernel, which returns dL/dw (CSR order,
through the normalisation) and dL/dt.  The adjoint reads the transition
weights resident in the graph, so when another forward replaced them
between this forward and its backward, backward loads the saved weights
again first.  Both run on the graph's stream, which is
ordered after torch's pending work (see _join_torch_stream).
"""
from __future__ import annotations

import torch

from . import Graph, _abi


def _join_torch_stream(graph: Graph):
    """Order the library's work after torch's pending work.  A real torch
    stream is shared outright; on the legacy default stream (handle 0) the
    library keeps its own non-blocking stream, so torch's work is drained
    first (the library synchronises its own stream before returning)."""
    handle = torch.cuda.current_stream().cuda_stream
    if handle:
        graph.set_stream(handle)
    else:
        torch.cuda.current_stream().synchronize()


class PageRankFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, weights, teleport, graph: Graph, damping: float, tol_l1: float, max_iter: int, adj_tol: float):
        if weights.dtype != torch.float32 or not weights.is_cuda or weights.numel() != graph.m:
            raise ValueError("weights must be a float32 CUDA tensor with one entry per CSR edge")
        if graph.value_dtype != _abi.GG_VALUES_F32:
            raise ValueError("the graph must be created with float32 edge values")
        w = weights.detach().contiguous()
        rank = torch.empty(graph.n, dtype=torch.float32, device=weights.device)
        tp = None if teleport is None else teleport.detach().contiguous()
        _join_torch_stream(graph)
        graph.update_values(w, device=True)
        graph.pagerank(damping=damping, teleport=tp, tol_l1=tol_l1, max_iter=max_iter, flags=_abi.GG_OUT_DEVICE,
                       out=rank)
        ctx.graph = graph
        ctx.damping = damping
        ctx.adj_tol = adj_tol
        ctx.max_iter = max_iter
        ctx.has_teleport = teleport is not None
        ctx.values_version = graph.values_version
        ctx.save_for_backward(rank, w)
        return rank

    @staticmethod
    def backward(ctx, grad_rank):
        rank, w = ctx.saved_tensors
        graph: Graph = ctx.graph
        c = grad_rank.detach().to(torch.float32).contiguous()
        gw = torch.empty(graph.m, dtype=torch.float32, device=rank.device)
        gt = torch.empty(graph.n, dtype=torch.float32, device=rank.device)
        _join_torch_stream(graph)
        if graph.values_version != ctx.values_version:
            # Another forward loaded other weights since this one ran.
            graph.update_values(w, device=True)
            ctx.values_version = graph.values_version
        graph.pagerank_adjoint(rank, c, damping=ctx.damping, tol_l1=ctx.adj_tol, max_iter=ctx.max_iter,
                               flags=_abi.GG_OUT_DEVICE, grad_w=gw, grad_t=gt)
        return gw, (gt if ctx.has_teleport else None), None, None, None, None, None


def pagerank(graph: Graph, weights: torch.Tensor, teleport: torch.Tensor | None = None, damping: float = 0.85,
             tol_l1: float = 2e-7, max_iter: int = 1000, adj_tol: float = 2e-7) -> torch.Tensor:
    """Ranks r(w, t) with gradients for the raw edge weights and teleport."""
    return PageRankFunction.apply(weights, teleport, graph, damping, tol_l1, max_iter, adj_tol)
