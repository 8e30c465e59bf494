"""A few calls of every solver path through the AddressSanitizer build of the
CPU emulation library (libgst_emu_asan.so, selected with GPU4GST_LIB and run
with libasan preloaded), as a child process of test_emu_asan.py: single plain
and hop-bounded queries, a banded batch, rolling admission with overflowing
worklists, tree recovery. Any out-of-bounds access of a kernel or of the host
code aborts the process with AddressSanitizer's report; costs must equal the
oracle's. Prints one JSON summary line: {"cases": n, "fails": [...]}.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "gpu4gst-sigmod-fullexp_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from gpu4gst import GSTGraph, workloads, STATUS_OK, STATUS_INFEASIBLE  # noqa: E402
import oracle as O  # noqa: E402


def check(g, grp, q, D, ref, status, cost, edges, tag, fails):
    if ref["cost"] >= O.INF():
        ok = status == STATUS_INFEASIBLE
    else:
        ok = status == STATUS_OK and cost == ref["cost"]
        if ok and len(edges):
            code, wsum = O.verify_tree(g, grp, q, edges, D)
            ok = code == 0 and wsum == cost
    if not ok:
        fails.append((tag, q.tolist(), D, cost, int(ref["cost"])))


def main():
    fails, cases = [], 0
    g, grp = workloads.small(60, 170, 8, seed_=91, whi=30, member_p=0.06)
    G = GSTGraph.from_generated(g)
    G.set_groups(grp)
    rng = np.random.default_rng(93)
    qs = [rng.choice(8, int(k), replace=False).astype(np.int32) for k in rng.integers(2, 8, 36)]
    Ds = rng.integers(1, 4, len(qs)).astype(np.int32)
    # single queries, plain and hop-bounded
    for i in range(3):
        for D in (-1, int(Ds[i])):
            r = G.solve(qs[i], D=D)
            ref = O.dpbf(g, grp, qs[i]) if D < 0 else O.hop(g, grp, qs[i], D)
            cases += 1
            check(g, grp, qs[i], D, ref, r.status, int(r.cost), r.edges, "single", fails)
    # one batch large enough for the A* bands, and its hop-bounded twin
    for D in (None, Ds):
        b = G.solve_batch(qs, D=D)
        for i, q in enumerate(qs):
            Dq = -1 if D is None else int(D[i])
            ref = O.dpbf(g, grp, q) if D is None else O.hop(g, grp, q, Dq)
            cases += 1
            check(g, grp, q, Dq, ref, b.status[i], int(b.costs[i]), b.tree(i), "batch", fails)
    G.close()
    # rolling admission under a resident cap with worklists far too small:
    # resident queries restart, the rest are solved again with full-size lists
    g, grp = workloads.small(120, 400, 7, seed_=95, whi=40, member_p=0.05)
    for D in (None, True):
        G = GSTGraph.from_generated(g)
        G.set_groups(grp)
        G.set_option("max_batch", 3)
        G.set_option("list_permille", 15)
        qs = [rng.choice(7, int(k), replace=False).astype(np.int32) for k in rng.integers(2, 6, 8)]
        Dv = rng.integers(1, 4, len(qs)).astype(np.int32) if D else None
        b = G.solve_batch(qs, D=Dv)
        for i, q in enumerate(qs):
            Dq = -1 if Dv is None else int(Dv[i])
            ref = O.dpbf(g, grp, q) if Dv is None else O.hop(g, grp, q, Dq)
            cases += 1
            check(g, grp, q, Dq, ref, b.status[i], int(b.costs[i]), b.tree(i), "rolling", fails)
        G.close()
    print(json.dumps({"cases": cases, "fails": fails}))


if __name__ == "__main__":
    main()
