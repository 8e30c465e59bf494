"""The reference-shaped C++ adapters over the gfx950 library.

tests/cpp/adapter_demo.cpp, built by build() against libgst.so, calls
gst_ref::dp_gpu / gst_ref::dp_gpu_hop (the call shapes of DP_GPU,
code/TrimCDP/include/DPQ.cuh:67, and DP_gpu, code/D-TrimCDP/include/
DPQ.cuh:66) and checks every returned tree with the reference's feasibility
report (gst_ref::feasible, after this_is_a_feasible_solution_gpu,
code/PrunedDP++/include/exp_CPU_nonHop.h:17-49), its edge-weight sum
(gst_ref::tree_cost) and PrunedDP++'s MST post-processing
(gst_ref::mst_postprocess, after
code/PrunedDP++/include/graph_hash_of_mixed_weighted_MST_postprocessing.h:9-34).
The same checks as the CPU emulation test, here on the GPU and at sizes up to
the 100-query R-MAT scale-15 workload.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from gpu4gst import workloads
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gpu4gst-sigmod-fullexp_amd", "_build", "adapter_demo_gpu")


def _exe():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built; run __graft_entry__.build()")
    return EXE


def test_adapter_selftest_gpu_build(gpu):
    r = subprocess.run([_exe(), "selftest"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = dict(line.split()[1:] for line in r.stdout.splitlines())
    assert got == {"path_feasible": "1", "path_cost": "7", "path_mst_cost": "3",
                   "split_feasible": "0", "split_why": "disconnected",
                   "part_feasible": "0", "part_why": "uncovered", "single_feasible": "1"}


def _run(g, grp, queries):
    lines = [f"{g.n} {g.m}"] + [f"{a} {b} {c}" for a, b, c in zip(g.src, g.dst, g.w)]
    lines.append(str(grp.count))
    for i in range(grp.count):
        m = grp.members(i)
        lines.append(" ".join(map(str, [len(m), *m])))
    lines.append(str(len(queries)))
    for D, q in queries:
        lines.append(" ".join(map(str, [D, len(q), *q])))
    print(f"adapter_demo n={g.n} m={g.m} queries={len(queries)}", flush=True)
    r = subprocess.run([_exe()], input="\n".join(lines) + "\n", capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    got = [json.loads(s) for s in r.stdout.strip().splitlines()]
    assert len(got) == len(queries)
    return got


def _check(g, grp, queries, got, oracle):
    n_feasible = 0
    for (D, q), res in zip(queries, got):
        qa = np.array(q, np.int32)
        ref = oracle(qa, D)
        if ref["status"] == O.INFEASIBLE:
            # the reference's conventions: 0 from DP_GPU, 100000 from DP_gpu
            assert res["status"] == 1 and res["cost"] == (0 if D < 0 else 100000)
            continue
        assert res["status"] == 0 and res["cost"] == ref["cost"], (q, D, res["cost"], ref["cost"])
        assert res["feasible"] == 1 and res["tree_cost"] == res["cost"]
        if D < 0:  # MST post-processing cannot improve an optimal plain tree
            assert res["mst_cost"] == res["cost"]
        else:  # nor make a hop-bounded optimum costlier
            assert res["mst_cost"] <= res["cost"]
        if res["edges"]:
            code, wsum = O.verify_tree(g, grp, qa, np.array(res["edges"], np.int32), D)
            assert code == 0 and wsum == res["cost"]
        n_feasible += 1
    return n_feasible


def test_adapters_small_graphs_gpu_build(gpu):
    for seed in range(3):
        g, grp = workloads.small(300, 900, 8, seed_=300 + seed, whi=40, member_p=0.02)
        rng = np.random.default_rng(seed)
        queries = []
        for _ in range(16):
            k = int(rng.integers(1, 8))
            q = rng.choice(8, k, replace=False).tolist()
            queries.append((int(rng.choice([-1, -1, 2, 4, 6])), q))
        got = _run(g, grp, queries)
        n = _check(g, grp, queries, got,
                   lambda qa, D: O.dpbf(g, grp, qa) if D < 0 else O.hop(g, grp, qa, D))
        assert n > 0


def test_adapters_rmat15_gpu_build(gpu):
    w = workloads.config(1, n_queries=30)
    queries = [(-1, q.tolist()) for q in w.queries[:24]] + [(3, q.tolist()) for q in
                                                           w.queries[24:30]]
    got = _run(w.graph, w.groups, queries)
    n = _check(w.graph, w.groups, queries, got,
               lambda qa, D: (O.prunedpp(w.graph, w.groups, qa) if D < 0
                              else O.hop(w.graph, w.groups, qa, D)))
    assert n >= 24
