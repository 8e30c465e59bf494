"""Randomised batched parity on the GPU: for each seed a random graph (30 to
1,500 vertices, weights from unit to 1,000), a batch of 40 plain queries of
1..7 groups and a batch of up to 30 hop-bounded ones (D = 0..5), under the
default settings, a resident cap of 3 (rolling admission), a fixed one-label
radius, or an HBM budget of a quarter of the footprint. Every cost must equal
the oracle's and every tree must verify.

Prints each seed's shape before its GPU calls.
"""
import numpy as np
import pytest

from gpu4gst import GSTGraph, workloads, STATUS_INFEASIBLE, STATUS_OK
import oracle as O

pytestmark = pytest.mark.gpu


def _ok(g, grp, q, D, status, cost, edges):
    ref = O.dpbf(g, grp, q) if D < 0 else O.hop(g, grp, q, D)
    if ref["status"] == O.INFEASIBLE:
        return status == STATUS_INFEASIBLE
    if not (status == STATUS_OK and cost == ref["cost"]):
        return False
    if len(edges):
        code, wsum = O.verify_tree(g, grp, q, edges, D)
        return code == 0 and wsum == cost
    return True


@pytest.mark.parametrize("block", range(4))
def test_random_batches_match_oracle(gpu, block):
    fails = []
    for seed in range(block * 16, block * 16 + 16):
        rng = np.random.default_rng(90000 + seed)
        n = int(rng.integers(30, 1500))
        m = int(rng.integers(n, 5 * n))
        ng = int(rng.integers(8, 14))
        g, grp = workloads.small(n, m, ng, seed_=9000 + seed,
                                 whi=int(rng.choice([1, 5, 60, 1000])),
                                 member_p=float(rng.choice([0.01, 0.03, 0.1])))
        G = GSTGraph.from_generated(g)
        G.set_groups(grp)
        qs = [rng.choice(ng, int(k), replace=False).astype(np.int32)
              for k in rng.integers(1, 8, 40)]
        mode = seed % 4
        if mode == 1:
            G.set_option("max_batch", 3)
        if mode == 2:
            G.set_option("radius_pct", int(rng.choice([50, 70, 100, 130])))
        budget = sum(G.footprint(len(q)) for q in qs) // 4 if mode == 3 else 0
        print(f"seed {seed} n={g.n} m={g.m} groups={ng} mode={mode}", flush=True)
        b = G.solve_batch(qs, hbm_budget=budget)
        for i, q in enumerate(qs):
            if not _ok(g, grp, q, -1, b.status[i], int(b.costs[i]), b.tree(i)):
                fails.append(("plain", seed, i, q.tolist(), int(b.costs[i])))
        qh = [q for q in qs if len(q) <= 5][:30]
        Ds = rng.integers(0, 6, len(qh)).astype(np.int32)
        bh = G.solve_batch(qh, D=Ds, hbm_budget=budget // 2)
        for i, q in enumerate(qh):
            if not _ok(g, grp, q, int(Ds[i]), bh.status[i], int(bh.costs[i]), bh.tree(i)):
                fails.append(("hop", seed, i, q.tolist(), int(Ds[i]), int(bh.costs[i])))
        G.close()
    assert fails == [], fails
