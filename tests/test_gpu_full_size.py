"""Full-size parity on the GPU against committed oracle costs.

tests/golden/full_size/*.json hold the restated PrunedDP++ / D-PrunedDP++
costs of the seeded workloads at their full sizes (written by
tests/golden/make_full_size_costs.py in the build container) together with
digests of the inputs. Each test regenerates its workload, checks the
digests, solves on the GPU and compares bit for bit; every returned tree is
verified (connected, acyclic, touching every query group, weighing the cost,
within the hop bound).

The differentiable workload (V = 317,000) also checks the backward: the
gradient of sum_q g_q cost_q is sum_q g_q x incidence(tree_q); for 16 queries
and 64 sampled edges each, single-query re-solves with w_e + 1 and w_e - 1
bracket the incidence; and where those re-solves show the optimum is unique
(raising any tree edge raises the cost) the tree equals PrunedDP++'s.
"""
import os
import sys

import numpy as np
import pytest
import torch

from gpu4gst import GSTGraph, workloads, STATUS_OK, STATUS_INFEASIBLE
from gpu4gst.autograd import gst_cost
import oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import full_size  # noqa: E402

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _load(cfg):
    w = workloads.config(cfg)
    fx = full_size.load(w.name)
    assert fx is not None, f"fixture for {w.name} missing"
    assert fx["digests"] == full_size.digests(w), "workload differs from the fixture's inputs"
    G = GSTGraph.from_generated(w.graph)
    G.set_groups(w.groups)
    return w, fx, G


def _verify(w, i, cost, edges, D=-1):
    q = w.queries[i]
    if len(edges) == 0:
        gs = np.zeros(w.graph.n, np.int64)
        for b, gid in enumerate(q):
            gs[w.groups.members(int(gid))] |= 1 << b
        assert cost == 0 and (gs == (1 << len(q)) - 1).any(), i
        return
    code, wsum = O.verify_tree(w.graph, w.groups, q, edges, D)
    assert code == 0 and wsum == cost, (i, code, wsum, cost)


def _compare(w, fx, res, D=None):
    n = 0
    for j, i in enumerate(fx["queries"]):
        st, cost = fx["status"][j], fx["costs"][j]
        if st == O.INFEASIBLE:
            assert res.status[i] == STATUS_INFEASIBLE, (i, int(res.status[i]))
            continue
        assert res.status[i] == STATUS_OK and int(res.costs[i]) == cost, (
            i, int(res.status[i]), int(res.costs[i]), cost)
        _verify(w, i, cost, res.tree(i), -1 if D is None else int(D[i]))
        n += 1
    return n


def test_golden_rmat15_all_queries(gpu):
    w, fx, G = _load(1)
    _say(f"{w.name}: {len(w.queries)} queries batched vs {len(fx['queries'])} golden costs")
    res = G.solve_batch(w.queries)
    assert len(fx["queries"]) == 100
    assert _compare(w, fx, res) == 100
    G.close()


def test_golden_chunglu168k_all_queries(gpu):
    w, fx, G = _load(2)
    _say(f"{w.name}: {len(w.queries)} queries batched vs {len(fx['queries'])} golden costs")
    res = G.solve_batch(w.queries)
    assert len(fx["queries"]) == 1024
    assert _compare(w, fx, res) == 1024
    for i in range(0, 1024, 128):
        _say(f"{w.name}: single query {i} k={len(w.queries[i])}")
        r = G.solve(w.queries[i])
        assert r.cost == res.costs[i]
        _verify(w, i, r.cost, r.edges)
    G.close()


@pytest.mark.timeout(300)
def test_golden_hop_1100k(gpu):
    w, fx, G = _load(3)
    _say(f"{w.name}: {len(w.queries)} hop queries batched (HBM admission) vs "
         f"{len(fx['queries'])} golden costs")
    res = G.solve_batch(w.queries, D=w.D)
    assert len(fx["queries"]) >= 64
    assert fx["D"] == [int(w.D[i]) for i in fx["queries"]]
    n = _compare(w, fx, res, w.D)
    _say(f"{w.name}: {n} feasible, {len(fx['queries']) - n} infeasible, "
         f"{res.stats['batches']} admission batches")
    for i in range(len(w.queries)):
        if res.status[i] == STATUS_OK:
            _verify(w, i, int(res.costs[i]), res.tree(i), int(w.D[i]))
    G.close()


def _tree_set(edges):
    return {(min(int(a), int(b)), max(int(a), int(b))) for a, b in np.asarray(edges).reshape(-1, 2)}


@pytest.mark.timeout(600)
def test_differentiable_317k_forward_backward(gpu):
    w, fx, G = _load(4)
    g = w.graph
    nq = len(w.queries)
    assert nq == 256 and len(fx["queries"]) == 256
    dev = torch.device("cuda:0")
    wt = torch.tensor(g.w.astype(np.float32), device=dev, requires_grad=True)
    _say(f"{w.name}: forward + backward of {nq} queries k=5, V={g.n}")
    kept = []
    costs = gst_cost(wt, G, w.queries, keep=kept)
    gout = torch.tensor(np.random.default_rng(4).uniform(0.5, 2.0, nq).astype(np.float32),
                        device=dev)
    (costs * gout).sum().backward()
    grad = wt.grad.detach().cpu().numpy().astype(np.float64)
    base = costs.detach().cpu().numpy().astype(np.int64)
    assert base.tolist() == fx["costs"], "forward costs differ from PrunedDP++"
    res = kept[0]  # the trees the backward scattered (ties may differ between calls)
    assert res.costs.astype(np.int64).tolist() == fx["costs"]
    go = gout.cpu().numpy().astype(np.float64)
    want = np.zeros(g.m)
    for q in range(nq):
        _verify(w, q, int(res.costs[q]), res.tree(q))
        np.add.at(want, res.tree_ids(q), go[q])
    assert np.allclose(grad, want, rtol=1e-5, atol=1e-4)

    # finite-difference bracket: 16 queries x 64 edges, single-query re-solves
    rng = np.random.default_rng(0x6753_5444)
    order = np.argsort(g.eid, kind="stable")  # CSR entries of undirected edge e: order[2e:2e+2]
    ent = order.reshape(-1, 2)
    adj_edges = [[] for _ in range(g.n)]
    for e in range(g.m):
        adj_edges[g.src[e]].append(e)
        adj_edges[g.dst[e]].append(e)
    checked = inside = unique = 0
    for q in range(16):
        c0 = int(res.costs[q])
        tree = np.unique(res.tree_ids(q))
        tsample = tree[:32]
        verts = np.unique(res.tree(q).reshape(-1))
        near = np.setdiff1d(np.unique(np.concatenate([adj_edges[v] for v in verts]))
                            if len(verts) else np.zeros(0, np.int64), tree)
        rest = 64 - len(tsample)
        pick_near = rng.choice(near, min(len(near), rest // 2), replace=False) if len(near) else []
        far = np.setdiff1d(rng.choice(g.m, rest, replace=False), np.concatenate([tree, near]))
        sample = np.concatenate([tsample, pick_near, far[:rest - len(pick_near)]]).astype(np.int64)
        _say(f"{w.name}: FD query {q} cost {c0}, tree {len(tree)} edges, {len(sample)} sampled")
        right1 = []
        for e in sample:
            inc = 1 if e in set(tree.tolist()) else 0
            we = int(g.w[e])
            G.update_weights(ent[e], we + 1)
            up = G.solve(w.queries[q]).cost
            right = up - c0
            if we > 1:
                G.update_weights(ent[e], we - 1)
                left = c0 - G.solve(w.queries[q]).cost
            else:
                left = 10 ** 9  # w_e - 1 = 0 is outside the weight domain
            G.update_weights(ent[e], we)
            checked += 1
            inside += int(right <= inc <= left)
            assert right <= inc <= left, (q, int(e), inc, right, left)
            if inc:
                right1.append(right == 1)
        if len(tsample) == len(tree) and all(right1):
            # every optimal tree holds every edge of this one: it is the only optimum
            unique += 1
            ref = _tree_set(fx["trees"][q])
            assert _tree_set(res.tree(q)) == ref, f"query {q}: unique optimum differs from PrunedDP++"
    _say(f"{w.name}: {inside}/{checked} sampled edges inside the bracket; "
         f"{unique} of 16 optima unique and equal to PrunedDP++'s tree")
    assert checked >= 16 * 60 and inside == checked
    assert unique >= 1
    assert G.solve_batch(w.queries[:16]).costs.astype(np.int64).tolist() == fx["costs"][:16]
    G.close()


@pytest.mark.timeout(900)
def test_rmat22_k7_hbm_admission(gpu):
    w, fx, G = _load(5)
    assert w.graph.n == 4194304
    nq = len(w.queries)
    _say(f"{w.name}: V={w.graph.n} E={w.graph.m}, {nq} queries k=7, "
         f"footprint {G.footprint(7) / 2**30:.2f} GiB per query")
    res = G.solve_batch(w.queries)
    _say(f"{w.name}: {res.stats['batches']} admission batches, {res.stats['reruns']} re-runs")
    assert res.stats["batches"] >= 2
    assert len(fx["queries"]) >= 4
    _compare(w, fx, res)
    for i in range(nq):
        assert res.status[i] == STATUS_OK
        _verify(w, i, int(res.costs[i]), res.tree(i))
    for i in (0, 97, 201):
        _say(f"{w.name}: single query {i}")
        assert G.solve(w.queries[i]).cost == res.costs[i]
    G.close()
