"""Differentiable optimal GST cost: torch.autograd.Function over libgst.so.

forward: exact int32 DP per query (weights must hold integer values >= 1);
backward: the hand-written gst_backward kernel scatters grad_out[q] onto the
undirected edges of query q's optimal tree, i.e. the edge-incidence
subgradient d cost_q / d w_e = 1[e in tree_q]. Cost is concave and
piecewise linear in each weight, so the incidence lies between the one-sided
finite differences cost(w + e) - cost(w) and cost(w) - cost(w - e).
The reference is not differentiable; this has no counterpart there.
"""
from __future__ import annotations

import numpy as np
import torch

from .solver import GSTGraph


class GSTCost(torch.autograd.Function):
    @staticmethod
    def forward(ctx, weights: torch.Tensor, graph: GSTGraph, queries, D=None, keep=None):
        if graph.eid is None:
            raise ValueError("graph needs undirected edge ids (GSTGraph.from_generated)")
        w = weights.detach().to("cpu", torch.float64).numpy()
        wi = np.rint(w).astype(np.int64)
        if not np.array_equal(wi.astype(np.float64), w) or wi.min() < 1:
            raise ValueError("weights must be integers >= 1 for the exact forward")
        graph.set_weights_csr(wi[graph.eid].astype(np.int32))
        res = graph.solve_batch(queries, D=D)
        dev = weights.device
        ctx.graph = graph
        ctx.nund = weights.shape[0]
        ctx.save_for_backward(torch.from_numpy(res.offsets).to(dev),
                              torch.from_numpy(res.edge_ids).to(dev))
        ctx.result = res
        if keep is not None:  # the caller's list receives the trees the backward uses
            keep.append(res)
        return torch.from_numpy(res.costs.astype(np.float32)).to(dev)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        off, ids = ctx.saved_tensors
        g = grad_out.detach().to(torch.float32).contiguous()
        grad_w = torch.zeros(ctx.nund, dtype=torch.float32, device=g.device)
        if g.is_cuda:
            torch.cuda.current_stream(g.device).synchronize()
            ctx.graph.backward(g.shape[0], off.data_ptr(), ids.data_ptr(), g.data_ptr(),
                               grad_w.data_ptr(), ctx.nund)
        else:
            raise RuntimeError("GSTCost.backward needs CUDA (HIP) tensors")
        return grad_w, None, None, None, None


def gst_cost(weights: torch.Tensor, graph: GSTGraph, queries, D=None, keep=None) -> torch.Tensor:
    """Optimal costs of `queries` as a float32 tensor [nq] differentiable in
    `weights` (float32 [num_undirected], integer values >= 1). When `keep` is a
    list, the forward's BatchResult (costs, statuses and the trees whose
    incidence the backward scatters) is appended to it: among tied optimal
    trees the solver may pick a different one on another call."""
    return GSTCost.apply(weights, graph, queries, D, keep)
