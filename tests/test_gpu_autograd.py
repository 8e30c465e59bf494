"""Backward of the optimal cost: edge-incidence subgradient inside the one-sided
finite-difference bracket cost(w+e) - cost(w) <= g_e <= cost(w) - cost(w-e)."""
import numpy as np
import pytest
import torch

from gpu4gst import GSTGraph, gen, workloads
from gpu4gst.autograd import gst_cost
import oracle as O

pytestmark = pytest.mark.gpu


def _fd_check(g, grp, queries, D, seed, n_edges=24):
    G = GSTGraph.from_generated(g)
    G.set_groups(grp)
    dev = torch.device("cuda:0")
    w = torch.tensor(g.w, dtype=torch.float32, device=dev, requires_grad=True)
    print(f"autograd n={g.n} m={g.m} nq={len(queries)} D={D} seed={seed}", flush=True)
    kept = []
    costs = gst_cost(w, G, queries, D, keep=kept)
    gout = torch.arange(1, len(queries) + 1, dtype=torch.float32, device=dev)
    (costs * gout).sum().backward()
    grad = w.grad.detach().cpu().numpy()
    base = costs.detach().cpu().numpy().astype(np.int64)
    # per-query incidence from the returned trees, checked against the oracle's costs
    res = kept[0]  # the trees the backward scattered
    inc = np.zeros((len(queries), g.m), np.int64)
    for q in range(len(queries)):
        inc[q, res.tree_ids(q)] = 1
        ref = (O.dpbf(g, grp, queries[q]) if D is None else O.hop(g, grp, queries[q], D))["cost"]
        assert base[q] == ref
    assert np.allclose(grad, (inc * np.arange(1, len(queries) + 1)[:, None]).sum(0))
    rng = np.random.default_rng(seed)
    tree_edges = np.nonzero(inc.any(0))[0]
    others = np.setdiff1d(np.arange(g.m), tree_edges)
    pick = np.concatenate([tree_edges[:n_edges // 2],
                           rng.choice(others, min(len(others), n_edges // 2), replace=False)])
    checked = 0
    for e in pick:
        for delta in (+1, -1):
            if delta < 0 and g.w[e] < 2:
                continue
            w2 = g.w.copy()
            w2[e] += delta
            G.set_weights_csr(w2[g.eid])
            c2 = G.solve_batch(queries, D=D).costs.astype(np.int64)
            for q in range(len(queries)):
                if delta > 0:
                    assert c2[q] - base[q] <= inc[q, e], (q, e)
                else:
                    assert inc[q, e] <= base[q] - c2[q], (q, e)
            checked += 1
    G.set_weights_csr(g.w[g.eid])
    assert checked > 0


@pytest.mark.parametrize("seed", range(3))
def test_fd_bracket_plain(gpu, seed):
    g, grp = workloads.small(80, 200, 8, seed_=40 + seed, whi=20, member_p=0.05)
    rng = np.random.default_rng(seed)
    qs = [rng.choice(8, int(k), replace=False).astype(np.int32) for k in (2, 3, 4, 5)]
    _fd_check(g, grp, qs, None, seed)


@pytest.mark.parametrize("seed", range(2))
def test_fd_bracket_hop(gpu, seed):
    g, grp = workloads.small(60, 150, 6, seed_=60 + seed, whi=15, member_p=0.06)
    rng = np.random.default_rng(10 + seed)
    qs = [rng.choice(6, int(k), replace=False).astype(np.int32) for k in (2, 3, 4)]
    _fd_check(g, grp, qs, 4, seed)
