"""Error and range handling of the solver, run through the CPU emulation build
(libgst_emu.so, selected with GPU4GST_LIB) as a child process of test_emu.py.
Prints one JSON summary line: {"cases": n, "fails": [...]}.

  * a tree longer than the recovery buffer (max_tree_edges = 1) is recovered
    by the retry with a larger buffer, not reported as an internal error;
  * a query whose only trees cost 2^29 or more is reported as out of range
    (status 2), not infeasible, while one just below 2^29 is solved exactly;
  * sparse weight updates: an asymmetric update is refused and changes
    nothing; a symmetric one gives the same costs as the oracle on the new
    weights;
  * the autotuner reports, tunes within its time budget, lets explicit
    options win and never changes a result;
  * trees from stored backpointers and from the cost table alone are both
    valid and optimal.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "gpu4gst-sigmod-fullexp_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from gpu4gst import (GSTError, GSTGraph, gen, workloads, STATUS_OK, STATUS_INFEASIBLE,  # noqa: E402
                     STATUS_RANGE)
import oracle as O  # noqa: E402


def path_graph(n, w):
    src = np.arange(n - 1, dtype=np.int32)
    g = gen.from_edges(n, src, src + 1, np.full(n - 1, w, np.int32))
    grp = gen.groups_from_lists([[0], [n - 1], [n // 2]])
    return g, grp


def main():
    fails, cases = [], 0

    def check(name, ok, detail=""):
        nonlocal cases
        cases += 1
        if not ok:
            fails.append((name, detail))

    # 1. long trees: buffer of one edge, trees of up to 39 edges
    g, grp = path_graph(40, 3)
    G = GSTGraph.from_generated(g)
    G.set_groups(grp)
    G.set_option("max_tree_edges", 1)
    gh, grph = path_graph(16, 2)
    Gh = GSTGraph.from_generated(gh)
    Gh.set_groups(grph)
    Gh.set_option("max_tree_edges", 1)
    for q, D in (([0, 1], 15), ([0, 1, 2], 15), ([1, 2], 9)):
        q = np.asarray(q, np.int32)
        r = Gh.solve(q, D=D)
        ref = O.hop(gh, grph, q, D)
        code, wsum = O.verify_tree(gh, grph, q, r.edges, D)
        check("long_hop_tree", r.status == STATUS_OK and r.cost == ref["cost"] and code == 0
              and wsum == r.cost, f"q={q.tolist()} D={D} cost={r.cost} ref={ref['cost']}")
    Gh.close()
    for q, D in (([0, 1], -1), ([0, 1, 2], -1), ([1, 2], -1)):
        q = np.asarray(q, np.int32)
        r = G.solve(q, D=D)
        ref = O.dpbf(g, grp, q) if D < 0 else O.hop(g, grp, q, D)
        code, wsum = O.verify_tree(g, grp, q, r.edges, D)
        check("long_tree", r.status == STATUS_OK and r.cost == ref["cost"] and code == 0
              and wsum == r.cost, f"q={q.tolist()} D={D} cost={r.cost} ref={ref['cost']}")
    b = G.solve_batch([np.asarray([0, 1], np.int32), np.asarray([1, 2], np.int32)])
    check("long_tree_batch", b.costs.tolist() == [117, 57] and len(b.tree(0)) == 39
          and len(b.tree(1)) == 19, f"{b.costs.tolist()}")
    G.close()

    # 2. the int32 cost range: a path of 600 edges of weight 2^20 costs
    # 629,145,600 >= 2^29; a path of 500 such edges costs 524,288,000 < 2^29.
    # (Hop-bounded trees cannot get there: with k <= 7 and D <= 15 an optimal
    # tree has at most 105 edges.)
    big = 1 << 20
    for n, want in ((601, STATUS_RANGE), (501, STATUS_OK)):
        g, grp = path_graph(n, big)
        G = GSTGraph.from_generated(g)
        G.set_groups(grp)
        r = G.solve(np.asarray([0, 1], np.int32))
        ok = r.status == want
        if want == STATUS_OK:
            ok = ok and r.cost == (n - 1) * big and len(r.edges) == n - 1
        check("range", ok, f"n={n} status={r.status} cost={r.cost}")
        G.close()
    # a hop-bounded query that is truly infeasible stays infeasible (no saturation)
    g, grp = path_graph(30, 5)
    G = GSTGraph.from_generated(g)
    G.set_groups(grp)
    r = G.solve(np.asarray([0, 1], np.int32), D=3)
    check("hop_infeasible", r.status == STATUS_INFEASIBLE, f"status={r.status}")
    G.close()

    # 3. sparse weight updates
    rng = np.random.default_rng(77)
    g, grp = workloads.small(40, 90, 6, seed_=77, whi=30, member_p=0.1)
    G = GSTGraph.from_generated(g)
    G.set_groups(grp)
    e = int(rng.integers(0, g.m))
    ent = np.nonzero(g.eid == e)[0]
    try:
        G.update_weights(ent[:1], [int(g.w[e]) + 5])
        check("asym_refused", False, "accepted")
    except GSTError:
        check("asym_refused", True)
    qs = [rng.choice(6, k, replace=False).astype(np.int32) for k in (2, 3, 4)]
    base = G.solve_batch(qs).costs
    check("unchanged_after_refusal",
          all(int(base[i]) == O.dpbf(g, grp, q)["cost"] for i, q in enumerate(qs)))
    for trial in range(6):
        es = rng.choice(g.m, 3, replace=False)
        w2 = g.w.copy()
        w2[es] = rng.integers(1, 40, 3)
        idx = np.concatenate([np.nonzero(g.eid == x)[0] for x in es])
        G.update_weights(idx, w2[g.eid[idx]])
        g2 = gen.Graph(g.n, g.src, g.dst, w2, g.beg, g.col, w2[g.eid].astype(np.int32), g.eid)
        got = G.solve_batch(qs).costs
        for i, q in enumerate(qs):
            ref = O.dpbf(g2, grp, q)["cost"]
            check("update", int(got[i]) == ref, f"trial={trial} q={q.tolist()} {got[i]} {ref}")
        G.update_weights(idx, g.w[g.eid[idx]])  # back to the original weights
    got = G.solve_batch(qs).costs
    check("restored", np.array_equal(got, base), f"{got.tolist()} {base.tolist()}")

    # 5. stored vs recomputed backpointers: same costs, valid trees either way
    for mode in (1, 0, -1):
        G.set_option("backpointers", mode)
        for D in (None, 3):
            r = G.solve_batch(qs, D=D)
            check("bp_flag", r.stats["backpointers_stored"] == (0 if mode == 0 else 1),
                  f"mode={mode} {r.stats['backpointers_stored']}")
            for i, q in enumerate(qs):
                ref = O.dpbf(g, grp, q) if D is None else O.hop(g, grp, q, D)
                if ref["status"] != 0:
                    check("bp_infeasible", r.status[i] == STATUS_INFEASIBLE)
                    continue
                if len(r.tree(i)) == 0:  # a vertex in every query group
                    check("bp_tree", int(r.costs[i]) == 0 == ref["cost"], f"q={q.tolist()}")
                    continue
                code, wsum = O.verify_tree(g, grp, q, r.tree(i), -1 if D is None else D)
                check("bp_tree", int(r.costs[i]) == ref["cost"] and code == 0
                      and wsum == ref["cost"], f"mode={mode} D={D} q={q.tolist()}")

    # 6. per-launch trace of a profiled call
    G.set_option("profile", 1)
    G.set_option("trace", 1)
    G.solve_batch(qs)
    rows = G.trace()
    G.set_option("trace", 0)
    G.set_option("profile", 0)
    check("trace", len(rows) > 2 and all(0 <= r["active_slots"] <= r["slots"] == len(qs)
                                          for r in rows)
          and {r["phase"] for r in rows} == {0, 1} and any(r["items"] > 0 for r in rows)
          and all(r["active_slots"] > 0 for r in rows if r["items"] > 0 and r["phase"] == 0),
          str(rows[:4]))

    # 4. autotuner: reports the class model without timing, tunes within its
    # budget, never changes results, refuses a bad batch size
    rep = G.autotune(qs, batch_size=1, max_seconds=0)
    check("tune_report", rep["size_class"] == 0 and rep["chunk"] == 256 and rep["trials"] == 0,
          str(rep))
    rep = G.autotune(qs[:1], batch_size=1, max_seconds=8.0)
    check("tune_single", rep["size_class"] == 0 and rep["trials"] >= 1
          and rep["seconds_after"] <= rep["seconds_before"], str(rep))
    rep = G.autotune(qs, batch_size=0, max_seconds=5.0)
    check("tune_batch", rep["size_class"] == 1 and rep["trials"] >= 1, str(rep))
    for i, q in enumerate(qs):
        check("tuned_single", G.solve(q).cost == base[i])
    check("tuned_batch", np.array_equal(G.solve_batch(qs).costs, base))
    G.set_option("chunk", 64)  # an explicit option wins over the tuned choice
    check("explicit_wins", G.autotune(qs, batch_size=1, max_seconds=0)["chunk"] == 64)
    G.set_option("chunk", 0)
    try:
        G.autotune(qs, batch_size=5)
        check("tune_bad_batch", False, "accepted")
    except GSTError:
        check("tune_bad_batch", True)
    G.close()
    # 7. the allocation-free entry point: two calls in a row on different
    # graphs, the second smaller, agree with gst_solve (the scratch array keeps
    # the first call's contents)
    import ctypes as C
    from gpu4gst import _lib
    L = _lib.lib()
    big, bgrp = workloads.small(120, 300, 8, seed_=11, whi=50, member_p=0.04)
    small, sgrp = workloads.small(30, 60, 4, seed_=12, whi=9, member_p=0.1)
    for gg, gp, q in ((big, bgrp, [0, 2, 4, 5, 7]), (small, sgrp, [1, 3])):
        Gx = GSTGraph.from_generated(gg)
        Gx.set_groups(gp)
        for D in (-1, 3):
            qa = np.asarray(q, np.int32)
            cost, ne, st = C.c_int32(0), C.c_int32(0), C.c_int32(0)
            uv = np.zeros(2 * 512, np.int32)
            rc = L.gst_solve_noalloc(Gx._h, qa.ctypes.data_as(C.c_void_p), len(q), D,
                                     C.byref(cost), uv.ctypes.data_as(C.c_void_p), 512,
                                     C.byref(ne), C.byref(st), None)
            ref = Gx.solve(qa, D=D)
            ok = rc == 0 and (cost.value, st.value) == (ref.cost, ref.status)
            if ok and st.value == 0 and ne.value:
                code, wsum = O.verify_tree(gg, gp, qa, uv[:2 * ne.value].reshape(-1, 2), D)
                ok = code == 0 and wsum == cost.value
            check("noalloc", ok, f"n={gg.n} q={q} D={D} rc={rc} {cost.value} {ref.cost}")
        Gx.close()

    print(json.dumps({"cases": cases, "fails": fails}))


if __name__ == "__main__":
    main()
