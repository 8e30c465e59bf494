"""Rolling admission through the CPU emulation build (libgst_emu.so, selected
with GPU4GST_LIB), as a child process of test_emu.py. Prints one JSON
summary line: {"cases": n, "rolling": n, "restarts": n, "fails": [...]}.

Calls whose queries do not all hold tables at once (a cap on resident
queries, or an HBM budget below the call's footprint) run under rolling
admission, plain and hop-bounded, with the DP unordered and in bands; every
cost must equal the oracle's and the lockstep batches' (GST_OPT_ROLLING = 0),
every tree must verify. Worklists far too small make the resident queries
restart, fewer at a time, with the same results.
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

DEFAULTS = {"max_batch": 0, "dp_delta_div": -1, "list_permille": 0, "rolling": 1}


def check(g, grp, qs, D, refs, res, tag, fails):
    for i, q in enumerate(qs):
        Dq = -1 if D is None else int(D[i])
        if refs[i] >= O.INF():
            ok = res.status[i] == STATUS_INFEASIBLE
        else:
            ok = res.status[i] == STATUS_OK and int(res.costs[i]) == refs[i]
            if ok and len(res.tree(i)):
                code, wsum = O.verify_tree(g, grp, q, res.tree(i), Dq)
                ok = code == 0 and wsum == refs[i]
        if not ok:
            fails.append((tag, i, q.tolist(), Dq, int(res.costs[i]), int(res.status[i]), refs[i]))


def main():
    fails, cases, rolled, restarts = [], 0, 0, 0
    for seed in range(2):
        rng = np.random.default_rng(7100 + seed)
        n = int(rng.integers(150, 260))
        g, grp = workloads.small(n, int(rng.integers(n, 3 * n)), 7, seed_=710 + seed, whi=40,
                                 member_p=0.04)
        for hop in (False, True):
            # a fresh graph: worklists only grow, so the restart case runs first
            G = GSTGraph.from_generated(g)
            G.set_groups(grp)
            qs = [rng.choice(7, int(k), replace=False).astype(np.int32)
                  for k in rng.integers(2, 7, 10)]
            D = np.array([int(x) for x in rng.integers(1, 4, len(qs))], np.int32) if hop else None
            refs = [(O.hop(g, grp, q, int(D[i])) if hop else O.dpbf(g, grp, q))["cost"]
                    for i, q in enumerate(qs)]
            foot = sum(G.footprint(len(q), -1 if D is None else int(D[i]))
                       for i, q in enumerate(qs))
            # lists of 1.5% of the states: the hop-bounded DP overflows them and
            # its resident queries restart; the plain calls overflow in the
            # one-label phase and are solved again with full-size lists
            runs = (("restart", dict(max_batch=4, list_permille=15), 0),
                    ("cap3", dict(max_batch=3), 0),
                    ("cap2-unordered", dict(max_batch=2, dp_delta_div=0), 0),
                    ("cap3-bands", dict(max_batch=3, dp_delta_div=8), 0),
                    ("budget", {}, foot // 3),
                    ("lockstep", dict(max_batch=3, rolling=0), 0))
            for tag, opts, budget in runs:
                for k, v in opts.items():
                    G.set_option(k, v)
                res = G.solve_batch(qs, D=D, hbm_budget=budget)
                cases += 1
                rolled += int(res.stats["rolling_queries"] > 0)
                restarts += int(res.stats["reruns"])
                check(g, grp, qs, D, refs, res, (seed, hop, tag), fails)
                for k in opts:
                    G.set_option(k, DEFAULTS[k])
            G.close()
    print(json.dumps({"cases": cases, "rolling": rolled, "restarts": restarts, "fails": fails}))


if __name__ == "__main__":
    main()
