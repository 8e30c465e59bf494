"""Full-size golden costs: where they live and how a workload is identified.

The fixtures under tests/golden/full_size/ hold the oracle's optimal costs
(and, for the differentiable workload, its trees) for the seeded workloads at
their full sizes, written by make_full_size_costs.py in the build container.
Each fixture records SHA-256 digests of the graph, the groups and the queries
it was computed on; a test regenerates the workload, checks the digests and
only then compares.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DIR = os.path.join(HERE, "full_size")


def digests(w) -> dict:
    g = hashlib.sha256()
    for a in (w.graph.beg, w.graph.col, w.graph.cw):
        g.update(np.ascontiguousarray(a).tobytes())
    gr = hashlib.sha256()
    gr.update(np.ascontiguousarray(w.groups.off).tobytes())
    gr.update(np.ascontiguousarray(w.groups.mem).tobytes())
    q = hashlib.sha256()
    for x in w.queries:
        q.update(np.asarray(len(x), np.int32).tobytes())
        q.update(np.ascontiguousarray(x, np.int32).tobytes())
    if w.D is not None:
        q.update(np.ascontiguousarray(w.D, np.int32).tobytes())
    return {"graph": g.hexdigest(), "groups": gr.hexdigest(), "queries": q.hexdigest(),
            "V": int(w.graph.n), "E_undirected": int(w.graph.m)}


def path(name: str) -> str:
    return os.path.join(DIR, f"{name}.json")


def load(name: str) -> dict | None:
    p = path(name)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)
