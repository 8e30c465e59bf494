"""Throughput benchmark of the batched TrimCDP solver on one MI355X.

Workload: Chung-Lu power-law graph, V = 168,000, 6.8M undirected edges
(exponent 2.1, largest expected degree 35,000), weights U[1,1000], 10,000
groups with Zipf(1.2) sizes in [1,5000], 1,024 queries with k cycling 3/5/7,
no hop bound. A step is one solve of all 1,024 queries in one batched call
(graph and groups resident in HBM before timing starts).

Printed JSON line: queries/s of the batched path (value), the single-query
path on a sample of the same queries, cost parity between the two, the
roofline of the dominant expand kernel (algorithmic bytes / measured kernel
time, from an instrumented pass timed with HIP events on the solver's stream)
and the CPU baselines on bounded samples: restated TrimCDP-multi-core-CPU on
the job's host threads (cpu_baseline) and restated PrunedDP++ on one thread
(cpu_baseline_one_core).

--workload 1/3/4/5 measures the other seeded workloads the same way
(3: hop-bounded, against restated HOP_MT / HOP_cpu; 4: forward + backward
through the autograd wrapper; 5: R-MAT scale 22, k = 7).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu4gst-sigmod-fullexp_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _traffic(workload: str) -> dict:
    """HBM bytes per launch of each expand kernel on this workload, from the
    committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE runs of this
    benchmark, each scaled by the reading of a copy kernel of known size)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get("kernels", {})


def _roof(kernel, bytes_alg, kernel_ms, launches, traffic, bytes_layout=None):
    """Roofline of one kernel: algorithmic bytes per launch / average launch
    time. For the DP expand kernel the bytes are 16 F + 12 Rg + 8 Rm + 12 U +
    4 Lb (frontier states, grow and merge relaxations, improvements, one-label
    terms of the bound tests); `achieved_layout` prices the same work with the
    bytes this build's layout moves (8-byte state words, 64-bit items, 32-byte
    one-label rows for upper bounds). `traffic_over_alg` = counter-measured HBM
    bytes per launch / algorithmic bytes per launch."""
    achieved = bytes_alg / (kernel_ms / 1e3) / 1e9 if kernel_ms > 0 else 0.0
    per_launch = bytes_alg / launches
    tr = traffic["bytes_per_launch"] if traffic else None
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": tr,
           "traffic_over_alg": tr / per_launch if tr and per_launch else None,
           "kernel": kernel, "kernel_ms_total": kernel_ms, "launches": launches,
           "bytes_alg_per_launch": per_launch,
           "kernel_ms_per_launch": kernel_ms / launches}
    if bytes_layout is not None and kernel_ms > 0:
        out["achieved_layout"] = bytes_layout / (kernel_ms / 1e3) / 1e9
    return out


def cpu_threads() -> int:
    """Host threads for the multi-core baselines: the CPU share the job was
    given (OMP_NUM_THREADS) or the visible cores, at most 100 as in the
    reference's pool (code/TrimCDP-multi-core-CPU/include/CPUNONHOP.h:692)."""
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    return max(1, min(n, 100))


def cpu_sample(w, budget_s: float, gpu, solve, name: str, cores: int):
    """Times `solve` (an oracle function) over the first queries of the
    workload until budget_s seconds have passed; costs (or infeasibility) are
    checked against the GPU's batch result `gpu`."""
    t0 = time.perf_counter()
    n, mismatches = 0, 0
    for i, q in enumerate(w.queries):
        r = solve(i, q)
        n += 1
        if i < len(gpu.costs):
            gpu_infeasible = int(gpu.status[i]) != 0
            cpu_infeasible = int(r["status"]) != 0
            if gpu_infeasible != cpu_infeasible or (
                    not gpu_infeasible and r["cost"] != int(gpu.costs[i])):
                mismatches += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    ks = sorted({len(q) for q in w.queries[:n]})
    return {"value": n / el, "unit": "queries/s", "cores": cores, "kind": "port",
            "sample": f"first {n} queries (k in {ks}) of the same workload, {name}, {el:.1f} s",
            "cost_mismatches_vs_gpu": mismatches}


def cpu_baselines(w, budget_s: float, gpu, D=None):
    """(multi-core baseline, one-core baseline): restated TrimCDP-multi-core-CPU
    cpu_mt and PrunedDP++ for plain queries, restated D-TrimCDP-multi-core-CPU
    HOP_MT and D-PrunedDP++ HOP_cpu for hop-bounded ones."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t = cpu_threads()
    if D is None:
        mt = cpu_sample(w, budget_s, gpu,
                        lambda i, q: O.trimcdp_mt(w.graph, w.groups, q, t),
                        f"restated TrimCDP-multi-core-CPU ({t} threads)", t)
        one = cpu_sample(w, budget_s, gpu, lambda i, q: O.prunedpp(w.graph, w.groups, q),
                         "restated PrunedDP++ (binary heap, array tables, 1 thread)", 1)
    else:
        mt = cpu_sample(w, budget_s, gpu,
                        lambda i, q: O.hop_mt(w.graph, w.groups, q, int(D[i]), t),
                        f"restated D-TrimCDP-multi-core-CPU HOP_MT ({t} threads)", t)
        one = cpu_sample(w, budget_s, gpu,
                         lambda i, q: O.hop(w.graph, w.groups, q, int(D[i])),
                         "restated D-PrunedDP++ HOP_cpu (1 thread)", 1)
    return mt, one


def _ablation(args) -> list:
    """Ablation switches in force (the reference's ablation variants)."""
    return [n for n, on in (("lane_appends", args.lane_appends),
                            ("per_item_edges", args.per_item_edges)) if on]


def _apply_options(G, args):
    if args.chunk:
        G.set_option("chunk", args.chunk)
    if args.sssp_delta >= 0:
        G.set_option("sssp_delta", args.sssp_delta)
    if args.dp_div >= 0:
        G.set_option("dp_delta_div", args.dp_div)
    if args.grid:
        G.set_option("grid", args.grid)
    if args.radius_pct:
        G.set_option("radius_pct", args.radius_pct)
    if args.max_batch:
        G.set_option("max_batch", args.max_batch)
    if args.sssp_slots >= 0:
        G.set_option("sssp_slots", args.sssp_slots)
    if args.lane_appends:
        G.set_option("lane_appends", 1)
    if args.per_item_edges:
        G.set_option("per_item_edges", 1)
    if args.rolling >= 0:
        G.set_option("rolling", args.rolling)


def _write_trace(G, path):
    """Per-launch trace of the profiled pass (phase, admission batch, items,
    queries that expanded, queries in the batch, ms) as CSV."""
    rows = G.trace()
    with open(path, "w") as f:
        f.write("launch,phase,batch,items,active_slots,slots,ms\n")
        for i, r in enumerate(rows):
            f.write(f"{i},{('one-label', 'dp', 'hop-dp')[r['phase']]},{r['batch']},{r['items']},"
                    f"{r['active_slots']},{r['slots']},{r['ms']:.4f}\n")
    log(f"trace: {len(rows)} launches -> {path}")


def _profiled_roofline(G, w, D, hop: bool, trace_path=None):
    """Instrumented solve: every expand launch timed with HIP events on the
    solver's stream; returns (dominant, other) roofline dicts."""
    G.set_option("profile", 1)
    if trace_path:
        G.set_option("trace", 1)
    pr = G.solve_batch(w.queries, D=D)
    G.set_option("profile", 0)
    if trace_path:
        _write_trace(G, trace_path)
        G.set_option("trace", 0)
    ps = pr.stats
    log(f"profiled: dp expand {ps['expand_ms']:.1f} ms / {ps['expand_launches']} launches, "
        f"one-label expand {ps['sssp_expand_ms']:.1f} ms")
    traffic = _traffic(w.name)
    # one-label expand: item 8 + row 32 + row bounds 8 B; edge visit: col 4 + weight 4 +
    # neighbour row 32 B; label update: 4 B
    sssp_bytes = 48 * ps["sssp_items"] + 40 * ps["sssp_relax"] + 4 * ps["sssp_updates"]
    roof_sssp = _roof("k_sssp_expand", sssp_bytes, ps["sssp_expand_ms"],
                      max(1, ps["sssp_expand_launches"]), traffic.get("k_sssp_expand"))
    kname = "k_dp_expand<true>" if hop else "k_dp_expand<false>"
    roof_dp = _roof(kname, ps["bytes_alg"], ps["expand_ms"], max(1, ps["expand_launches"]),
                    traffic.get(kname), ps["bytes_layout"])
    if ps["sssp_expand_ms"] >= ps["expand_ms"]:
        return roof_sssp, roof_dp
    return roof_dp, roof_sssp


def _dp_stats(st):
    return {k: st[k] for k in ("sssp_iterations", "sssp_items", "sssp_relax", "sssp_updates",
                               "sssp_items_unbounded", "dp_iterations", "states_expanded",
                               "grow_relax", "merge_relax", "improvements", "max_frontier",
                               "batches", "reruns", "rolling_queries", "max_resident")}


def other_workload(args):
    """Workloads 1, 3, 4, 5: batched throughput (hop-bounded for 3, forward +
    backward through the autograd wrapper for 4), single-query sample, cost
    parity, roofline of the dominant kernel and a CPU baseline sample."""
    import torch
    from gpu4gst import GSTGraph, workloads
    t0 = time.perf_counter()
    w = workloads.config(args.workload, n_queries=args.queries or None)
    log(f"workload {w.name}: V={w.graph.n} E={w.graph.m} groups={w.groups.count} "
        f"queries={len(w.queries)} ({time.perf_counter() - t0:.1f} s to generate)")
    G = GSTGraph.from_generated(w.graph)
    G.set_groups(w.groups)
    _apply_options(G, args)
    D = None if w.D is None else w.D
    hop = D is not None
    torch.cuda.synchronize()
    if args.profile_only:
        G.set_option("profile", 1)
        ps = G.solve_batch(w.queries, D=D).stats
        print(json.dumps({"profile_only": w.name, "expand_ms": ps["expand_ms"],
                          "expand_launches": ps["expand_launches"],
                          "sssp_expand_ms": ps["sssp_expand_ms"],
                          "sssp_expand_launches": ps["sssp_expand_launches"],
                          "phase_ms": ps["phase_ms"], "batches": ps["batches"]}))
        return
    grad = args.workload == 4
    if grad:
        from gpu4gst.autograd import gst_cost
        wt = torch.tensor(w.graph.w.astype(np.float32), device="cuda:0", requires_grad=True)

    def step():
        if grad:
            c = gst_cost(wt, G, w.queries)
            wt.grad = None
            c.sum().backward()
            return c.detach().cpu().numpy().astype(np.int64)
        return G.solve_batch(w.queries, D=D).costs

    for i in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    costs = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    qps = args.steps * len(w.queries) / el
    base = G.solve_batch(w.queries, D=D)
    stable = all(np.array_equal(np.asarray(c, np.int64), base.costs.astype(np.int64))
                 for c in costs)
    st = base.stats
    log(f"batched: {qps:.1f} queries/s, {el / args.steps * 1e3:.1f} ms/step")
    ns = min(args.single_sample, len(w.queries))
    if args.single_chunk:
        G.set_option("chunk", args.single_chunk)
    t1 = time.perf_counter()
    single = {}
    for i in range(ns):
        single[i] = G.solve(w.queries[i], D=-1 if D is None else int(D[i])).cost
    el1 = time.perf_counter() - t1
    mism = sum(int(base.costs[i] != single[i]) for i in range(ns))
    G.set_option("chunk", args.chunk)  # 0 = automatic again
    dominant, other = _profiled_roofline(G, w, D, hop, args.trace)
    out = {
        "metric": ("GST queries/s, forward + backward (batched, exact cost)" if grad else
                   "GST queries/s (batched, exact cost" + (", hop-bounded)" if hop else ")")),
        "value": qps, "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded generators)",
        "config": {"workload": w.name, "V": w.graph.n, "E_undirected": w.graph.m,
                   "groups": w.groups.count, "queries_per_step": len(w.queries),
                   "k": sorted({len(q) for q in w.queries}),
                   "hop_bound": None if D is None else sorted({int(x) for x in D}),
                   "ablation": _ablation(args)},
        "phase_ms": st["phase_ms"], "single_query_qps": ns / el1, "single_sample": ns,
        "batched_over_single": qps / (ns / el1),
        "cost_mismatch_batched_vs_single": mism, "costs_stable_across_steps": stable,
        "dp_stats": _dp_stats(st), "roofline": dominant, "roofline_other_kernel": other,
    }
    if not args.no_cpu:
        out["cpu_baseline"], out["cpu_baseline_one_core"] = cpu_baselines(
            w, args.cpu_seconds, base, D)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--queries", type=int, default=0,
                    help="queries per step (0 = the workload's own count)")
    ap.add_argument("--single-sample", type=int, default=96)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="time budget of each CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--sssp-delta", type=int, default=-1)
    ap.add_argument("--dp-div", type=int, default=-1,
                    help="A* band divisor of the plain DP (0 = unordered, -1 = automatic)")
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--lane-appends", action="store_true",
                    help="ablation: per-lane global appends (no LDS aggregation)")
    ap.add_argument("--per-item-edges", action="store_true",
                    help="ablation: each lane walks its own item's edges (no balanced mapping)")
    ap.add_argument("--radius-pct", type=int, default=0)
    ap.add_argument("--max-batch", type=int, default=0,
                    help="cap on queries per lockstep batch (0 = as many as the HBM budget admits)")
    ap.add_argument("--sssp-slots", type=int, default=-1,
                    help="queries per one-label pass inside a batch (0 = all at once)")
    ap.add_argument("--single-chunk", type=int, default=0,
                    help="virtual-split chunk for the single-query path (0 = the library's "
                         "automatic choice for one-query calls)")
    ap.add_argument("--profile-only", action="store_true",
                    help="only run the instrumented solve (the pass rocprofv3 summaries cover)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device copy of this size (counter calibration) and exit")
    ap.add_argument("--single-dp-div", type=int, default=-1)
    ap.add_argument("--rolling", type=int, default=-1,
                    help="1 = rolling admission when a call's tables do not fit at once (the "
                         "library default), 0 = lockstep admission batches")
    ap.add_argument("--trace", default="",
                    help="write the per-launch trace of the profiled pass to this CSV file")
    ap.add_argument("--workload", type=int, default=2,
                    help="2 = the headline workload (default); 1, 3, 4, 5 = the other seeded "
                         "workloads (3 hop-bounded, 4 forward + backward, 5 R-MAT scale 22)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (queries are "
                                   "independent: a batch shards across GPUs as one process "
                                   "per GPU with no exchange)"}))
        sys.exit(2)

    import torch
    if args.copy_probe_mib:
        # reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE: reads and
        # writes exactly copy_probe_mib MiB
        n = args.copy_probe_mib << 18
        a = torch.ones(n, dtype=torch.int32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_mib": args.copy_probe_mib}))
        return
    if args.workload != 2:
        other_workload(args)
        return
    from gpu4gst import GSTGraph, workloads
    t0 = time.perf_counter()
    w = workloads.config(2, n_queries=args.queries)
    log(f"workload {w.name}: V={w.graph.n} E={w.graph.m} groups={w.groups.count} "
        f"queries={len(w.queries)} ({time.perf_counter() - t0:.1f} s to generate)")
    assert torch.cuda.is_available()
    G = GSTGraph.from_generated(w.graph)
    G.set_groups(w.groups)
    _apply_options(G, args)
    torch.cuda.synchronize()

    if args.profile_only:
        G.set_option("profile", 1)
        pr = G.solve_batch(w.queries)
        ps = pr.stats
        print(json.dumps({"profile_only": w.name, "expand_ms": ps["expand_ms"],
                          "expand_launches": ps["expand_launches"],
                          "sssp_expand_ms": ps["sssp_expand_ms"],
                          "sssp_expand_launches": ps["sssp_expand_launches"]}))
        return
    for i in range(args.warmup):
        r = G.solve_batch(w.queries)
        log(f"warmup {i}: {r.stats['seconds_total']:.3f} s")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    results = []
    for i in range(args.steps):
        results.append(G.solve_batch(w.queries))
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_step = el / args.steps * 1e3
    qps = args.steps * len(w.queries) / el
    base = results[0]
    stable = all(np.array_equal(r.costs, base.costs) for r in results)
    st = base.stats
    log(f"batched: {qps:.1f} queries/s, {ms_step:.1f} ms/step, stats {st}")

    # single-query path on a sample (one query per call, same GPU), with its own tuning
    ns = min(args.single_sample, len(w.queries))
    if args.single_chunk:
        G.set_option("chunk", args.single_chunk)
    if args.single_dp_div >= 0:
        G.set_option("dp_delta_div", args.single_dp_div)
    G.solve(w.queries[0])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    single = {}
    for i in range(ns):
        single[i] = G.solve(w.queries[i]).cost
    torch.cuda.synchronize()
    el1 = time.perf_counter() - t1
    single_qps = ns / el1
    mism_single = sum(int(base.costs[i] != single[i]) for i in range(ns))
    log(f"single: {single_qps:.1f} queries/s over {ns}; mismatches {mism_single}")

    if args.single_chunk:
        G.set_option("chunk", args.chunk)  # 0 = automatic again
    if args.single_dp_div >= 0:
        G.set_option("dp_delta_div", args.dp_div)  # -1 = automatic
    dominant, other = _profiled_roofline(G, w, None, False, args.trace)
    relax = st["grow_relax"] + st["merge_relax"]

    out = {
        "metric": "GST queries/s at k=3/5/7 (batched, exact cost)",
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded Chung-Lu graph, Zipf groups, uniform queries)",
        "config": {"workload": w.name, "V": w.graph.n, "E_undirected": w.graph.m,
                   "groups": w.groups.count, "queries_per_step": len(w.queries),
                   "k": [3, 5, 7], "hop_bound": None, "ablation": _ablation(args)},
        "phase_ms": st["phase_ms"],
        "single_query_qps": single_qps,
        "single_query_chunk": args.single_chunk or "automatic",
        "batched_over_single": qps / single_qps if single_qps else None,
        "single_sample": ns,
        "cost_mismatch_batched_vs_single": mism_single,
        "costs_stable_across_steps": stable,
        "state_relaxations_per_s": relax * args.steps / el,
        "dp_stats": _dp_stats(st),
        "roofline": dominant,
        "roofline_other_kernel": other,
    }
    if not args.no_cpu:
        out["cpu_baseline"], out["cpu_baseline_one_core"] = cpu_baselines(
            w, args.cpu_seconds, base)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
