"""Runs small TrimCDP / D-TrimCDP cases through the CPU emulation build of the
kernels (libgst_emu.so, selected with GPU4GST_LIB) and compares them with the
oracle. Started as a child process by test_emu.py so that the emulation
library is the only libgst loaded in that process. Prints one JSON summary.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "gpu4gst-sigmod-fullexp_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from gpu4gst import GSTError, GSTGraph, gen, workloads, STATUS_OK, STATUS_INFEASIBLE  # noqa: E402
import oracle as O  # noqa: E402


def check_tree(g, grp, q, cost, edges, D=-1):
    if len(edges) == 0:
        gs = np.zeros(g.n, np.int64)
        for i, gid in enumerate(q):
            gs[grp.members(int(gid))] |= 1 << i
        return cost == 0 and bool((gs == (1 << len(q)) - 1).any())
    code, wsum = O.verify_tree(g, grp, q, edges, D)
    return code == 0 and wsum == cost


def main(n_plain, n_hop):
    fails, cases = [], 0
    for seed in range(n_plain):
        rng = np.random.default_rng(5000 + seed)
        n = int(rng.integers(4, 40))
        g, grp = workloads.small(n, int(rng.integers(0, 3 * n)), 6, seed_=500 + seed,
                                 whi=int(rng.choice([1, 7, 60])),
                                 member_p=float(rng.choice([0.05, 0.2])))
        G = GSTGraph.from_generated(g)
        G.set_groups(grp)
        qs = [rng.choice(6, k, replace=False).astype(np.int32) for k in (1, 3, 5, 6)]
        for q in qs:
            cases += 1
            r = G.solve(q)
            ref = O.dpbf(g, grp, q)
            if ref["cost"] >= O.INF():
                ok = r.status == STATUS_INFEASIBLE
            else:
                ok = r.status == STATUS_OK and r.cost == ref["cost"] and check_tree(
                    g, grp, q, r.cost, r.edges)
            if not ok:
                fails.append(("plain", seed, q.tolist(), int(r.cost), int(ref["cost"])))
        # the same queries in one lockstep batch, with the one-label phase in one
        # pass and in passes of two queries
        for per_pass in (0, 2):
            G.set_option("sssp_slots", per_pass)
            b = G.solve_batch(qs)
            for i, q in enumerate(qs):
                ref = O.dpbf(g, grp, q)
                if ref["cost"] < O.INF() and int(b.costs[i]) != ref["cost"]:
                    fails.append(("batch", per_pass, seed, q.tolist(), int(b.costs[i]),
                                  int(ref["cost"])))
        G.set_option("sssp_slots", 0)
        # single queries again after the batch: the slot layout changes between
        # calls, so stamps left by the batch sit where these queries keep theirs
        for q in reversed(qs):
            r = G.solve(q)
            ref = O.dpbf(g, grp, q)
            if ref["cost"] < O.INF() and r.cost != ref["cost"]:
                fails.append(("after-batch", seed, q.tolist(), int(r.cost), int(ref["cost"])))
        G.close()
    for seed in range(n_hop):
        rng = np.random.default_rng(6000 + seed)
        n = int(rng.integers(4, 30))
        g, grp = workloads.small(n, int(rng.integers(0, 2 * n)), 5, seed_=600 + seed,
                                 whi=int(rng.choice([1, 9])), member_p=0.2)
        G = GSTGraph.from_generated(g)
        G.set_groups(grp)
        for D in (0, 1, 3):
            q = rng.choice(5, int(rng.integers(1, 5)), replace=False).astype(np.int32)
            cases += 1
            r = G.solve(q, D=D)
            ref = O.hop(g, grp, q, D)
            if ref["cost"] >= O.INF():
                ok = r.status == STATUS_INFEASIBLE
            else:
                ok = r.status == STATUS_OK and r.cost == ref["cost"] and check_tree(
                    g, grp, q, r.cost, r.edges, D)
            if not ok:
                fails.append(("hop", seed, D, q.tolist(), int(r.cost), int(ref["cost"])))
        G.close()
    # calls of different shapes on one graph: a batch of mixed k, then the same
    # queries one at a time (each call lays out its slots differently)
    g, grp = workloads.small(150, 400, 10, seed_=17, whi=60, member_p=0.04)
    G = GSTGraph.from_generated(g)
    G.set_groups(grp)
    rng = np.random.default_rng(21)
    qs = [rng.choice(10, int(k), replace=False).astype(np.int32) for k in rng.integers(2, 8, 24)]
    refs = [O.dpbf(g, grp, q)["cost"] for q in qs]
    br = G.solve_batch(qs)
    for i in reversed(range(len(qs))):
        r = G.solve(qs[i])
        cases += 1
        for got, what in ((int(br.costs[i]), "mixed-batch"), (int(r.cost), "mixed-single")):
            if refs[i] < O.INF() and got != refs[i]:
                fails.append((what, i, qs[i].tolist(), got, int(refs[i])))
    G.close()
    # a batch large enough for the automatic A* bands (near/far lists)
    g, grp = workloads.small(80, 220, 10, seed_=23, whi=30, member_p=0.05)
    G = GSTGraph.from_generated(g)
    G.set_groups(grp)
    rng = np.random.default_rng(29)
    qs = [rng.choice(10, int(k), replace=False).astype(np.int32) for k in rng.integers(2, 8, 36)]
    br = G.solve_batch(qs)
    for i, q in enumerate(qs):
        cases += 1
        ref = O.dpbf(g, grp, q)["cost"]
        if ref < O.INF() and int(br.costs[i]) != ref:
            fails.append(("banded-batch", i, q.tolist(), int(br.costs[i]), int(ref)))
    G.close()
    # the ablation switches (per-lane appends, per-item edge walks, small
    # chunks): same results, plain and hop-bounded
    g, grp = workloads.small(70, 190, 8, seed_=37, whi=40, member_p=0.05)
    G = GSTGraph.from_generated(g)
    G.set_groups(grp)
    G.set_option("lane_appends", 1)
    G.set_option("per_item_edges", 1)
    G.set_option("chunk", 64)
    rng = np.random.default_rng(41)
    qs = [rng.choice(8, int(k), replace=False).astype(np.int32) for k in rng.integers(2, 7, 10)]
    Ds = rng.integers(1, 5, 10).astype(np.int32)
    bp = G.solve_batch(qs)
    bh = G.solve_batch(qs, D=Ds)
    for i, q in enumerate(qs):
        for res, ref, what in ((bp, O.dpbf(g, grp, q), "ablation-plain"),
                               (bh, O.hop(g, grp, q, int(Ds[i])), "ablation-hop")):
            cases += 1
            ok = ref["status"] == O.OK
            if ok != (int(res.status[i]) == STATUS_OK) or (ok and int(res.costs[i]) != ref["cost"]):
                fails.append((what, i, q.tolist(), int(res.costs[i]), int(ref["cost"])))
    G.close()
    # edge cases: an empty group, the largest hop bound, k = 7 hop-bounded, and
    # queries the solver must refuse (a group named twice, k = 8, D = 16, an
    # unknown group id, k = 0)
    g, grp = workloads.small(30, 70, 6, seed_=43, whi=20, member_p=0.2)
    lists = [grp.members(i).tolist() for i in range(grp.count)] + [[]]
    grp = gen.groups_from_lists(lists)
    empty = grp.count - 1
    full = [i for i in range(empty) if len(lists[i])]
    G = GSTGraph.from_generated(g)
    G.set_groups(grp)
    edge = [("empty-group", np.array([full[0], empty], np.int32), -1),
            ("empty-group-hop", np.array([empty, full[1]], np.int32), 2),
            ("max-hop-bound", np.array(full[:6], np.int32), 15),
            ("k7-hop", np.array(full[:7], np.int32), 3)]
    for what, q, D in edge:
        cases += 1
        r = G.solve(q, D=D)
        ref = O.dpbf(g, grp, q) if D < 0 else O.hop(g, grp, q, D)
        ok = ref["status"] == O.OK
        if ok != (r.status == STATUS_OK) or (ok and r.cost != ref["cost"]):
            fails.append((what, q.tolist(), int(r.cost), int(ref["cost"])))
        elif ok and not check_tree(g, grp, q, r.cost, r.edges, D):
            fails.append((what + "-tree", q.tolist(), int(r.cost), int(ref["cost"])))
    for what, q, D in (("repeated-group", np.array([full[0], full[1], full[0]], np.int32), -1),
                       ("k8", np.arange(8, dtype=np.int32), -1),
                       ("hop-bound-16", np.array(full[:2], np.int32), 16),
                       ("group-id-range", np.array([grp.count], np.int32), -1),
                       ("k0", np.zeros(0, np.int32), -1)):
        cases += 1
        try:
            G.solve(q, D=D)
            fails.append(("accepted " + what, q.tolist(), 0, 0))
        except GSTError as e:
            if e.code != 1:  # GST_ERR_INVALID
                fails.append(("error code " + what, q.tolist(), e.code, 1))
    G.close()
    print(json.dumps({"cases": cases, "fails": fails}))


if __name__ == "__main__":
    main(int(sys.argv[1]), int(sys.argv[2]))
