"""Writes tests/golden/full_size/<workload>.json: optimal costs from the CPU
oracle (oracle/liboracle.so, the restated PrunedDP++ and D-PrunedDP++) for
the seeded workloads at full size, with digests of the exact inputs.

Run in the build container (no GPU needed), from the repository root:
    python tests/golden/make_full_size_costs.py [workload ...]
Default: every workload. Queries are solved in parallel worker processes,
one query per worker at a time (each worker is single-threaded).

  1  all 100 queries, PrunedDP++
  2  all 1,024 queries, PrunedDP++
  3  the first 96 queries, D-PrunedDP++ HOP_cpu with each query's hop bound
  4  all 256 queries, PrunedDP++ with its trees (edge-incidence references)
  5  the first 6 queries, PrunedDP++ (about 2 minutes and 12 GB each)
"""
from __future__ import annotations

import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "gpu4gst-sigmod-fullexp_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import full_size  # noqa: E402

PLAN = {  # workload -> (number of queries, oracle, workers, keep trees)
    1: (None, "prunedpp", 8, False),
    2: (None, "prunedpp", 8, False),
    3: (96, "hop", 6, False),
    4: (None, "prunedpp", 8, True),
    5: (6, "prunedpp", 3, False),
}

_W = None


def _solve(i):
    import oracle as O
    w = _W
    t = time.perf_counter()
    if w.D is None:
        r = O.prunedpp(w.graph, w.groups, w.queries[i])
    else:
        r = O.hop(w.graph, w.groups, w.queries[i], int(w.D[i]))
    el = time.perf_counter() - t
    code = -1
    if r["status"] == 0:
        code, wsum = O.verify_tree(w.graph, w.groups, w.queries[i], r["edges"],
                                   D=-1 if w.D is None else int(w.D[i]))
        assert code == 0 and wsum == r["cost"], (i, code, wsum, r["cost"])
    return i, int(r["status"]), int(r["cost"]), r["edges"].tolist(), el


def make(cfg: int) -> None:
    global _W
    from gpu4gst import workloads
    n, oracle_name, workers, trees = PLAN[cfg]
    w = workloads.config(cfg)
    _W = w
    nq = len(w.queries) if n is None else min(n, len(w.queries))
    t0 = time.perf_counter()
    out = {}
    with mp.get_context("fork").Pool(workers) as pool:
        for i, st, cost, edges, el in pool.imap_unordered(_solve, range(nq)):
            out[i] = (st, cost, edges)
            print(f"  {w.name} query {i}: status {st} cost {cost} ({el:.1f} s)", flush=True)
    idx = list(range(nq))
    doc = {
        "workload": w.name,
        "config": cfg,
        "digests": full_size.digests(w),
        "oracle": {"prunedpp": "restated PrunedDP++ (oracle/gst_oracle.cpp orc_prunedpp)",
                   "hop": "restated D-PrunedDP++ HOP_cpu (oracle/gst_oracle.cpp orc_hop)"}[
                       oracle_name],
        "queries": idx,
        "status": [out[i][0] for i in idx],
        "costs": [out[i][1] for i in idx],
        "D": None if w.D is None else [int(w.D[i]) for i in idx],
        "seconds": round(time.perf_counter() - t0, 1),
    }
    if trees:
        doc["trees"] = [out[i][2] for i in idx]
    os.makedirs(full_size.DIR, exist_ok=True)
    with open(full_size.path(w.name), "w") as f:
        json.dump(doc, f, separators=(",", ":"))
        f.write("\n")
    print(f"{w.name}: {nq} queries in {doc['seconds']} s -> {full_size.path(w.name)}", flush=True)


if __name__ == "__main__":
    cfgs = [int(x) for x in sys.argv[1:]] or [1, 4, 2, 3, 5]
    for c in cfgs:
        make(c)
