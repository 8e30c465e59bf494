"""GPU parity: libgst.so against the CPU oracle, bit-exact costs, verified trees.

Each case prints its name, seed and shape before the GPU call.
"""
import numpy as np
import pytest

from gpu4gst import GSTGraph, gen, workloads, STATUS_INFEASIBLE, STATUS_OK
import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _graph(g, grp):
    G = GSTGraph.from_generated(g)
    G.set_groups(grp)
    return G


def _check_tree(g, grp, q, res_cost, edges, D=-1):
    if len(edges) == 0:
        # a single vertex in every query group
        gs = np.zeros(g.n, np.int64)
        for i, gid in enumerate(q):
            gs[grp.members(int(gid))] |= 1 << i
        assert res_cost == 0 and (gs == (1 << len(q)) - 1).any()
        return
    code, wsum = O.verify_tree(g, grp, q, edges, D)
    assert code == 0, f"tree check failed with code {code}"
    assert wsum == res_cost


@pytest.mark.parametrize("seed", range(24))
def test_plain_random_small(gpu, seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(4, 90))
    m = int(rng.integers(0, 3 * n))
    g, grp = workloads.small(n, m, 8, seed_=seed, whi=int(rng.choice([1, 3, 20, 100])),
                             member_p=float(rng.choice([0.03, 0.1, 0.3])))
    G = _graph(g, grp)
    for k in (1, 2, 3, 4, 5, 6, 7):
        q = rng.choice(8, k, replace=False).astype(np.int32)
        _say(f"plain seed={seed} n={g.n} m={g.m} k={k} q={q.tolist()}")
        r = G.solve(q)
        ref = O.dpbf(g, grp, q)
        assert r.status == STATUS_OK
        assert r.cost == ref["cost"], (r.cost, ref["cost"])
        _check_tree(g, grp, q, r.cost, r.edges)


@pytest.mark.parametrize("seed", range(12))
def test_hop_random_small(gpu, seed):
    rng = np.random.default_rng(2000 + seed)
    n = int(rng.integers(4, 60))
    m = int(rng.integers(0, 2 * n))
    g, grp = workloads.small(n, m, 6, seed_=100 + seed, whi=int(rng.choice([1, 5, 50])),
                             member_p=float(rng.choice([0.05, 0.15])))
    G = _graph(g, grp)
    for k in (1, 2, 3, 4, 5):
        q = rng.choice(6, k, replace=False).astype(np.int32)
        for D in (0, 1, 2, 3, 5):
            _say(f"hop seed={seed} n={g.n} m={g.m} k={k} D={D} q={q.tolist()}")
            r = G.solve(q, D=D)
            ref = O.hop(g, grp, q, D)
            if ref["status"] == O.INFEASIBLE:
                assert r.status == STATUS_INFEASIBLE
                continue
            assert r.cost == ref["cost"], (r.cost, ref["cost"])
            _check_tree(g, grp, q, r.cost, r.edges, D)


def test_hop_prune_off_matches(gpu):
    g, grp = workloads.small(50, 90, 6, seed_=5, whi=40)
    G = _graph(g, grp)
    q = np.array([0, 1, 2, 3], np.int32)
    _say("hop prune on/off n=50 k=4 D=3")
    a = G.solve(q, D=3).cost
    G.set_option("hop_prune", 0)
    b = G.solve(q, D=3).cost
    assert a == b == O.hop(g, grp, q, 3)["cost"]


def test_batch_equals_single(gpu):
    g, grp = workloads.small(120, 300, 10, seed_=11, whi=50, member_p=0.05)
    G = _graph(g, grp)
    rng = np.random.default_rng(3)
    qs = [rng.choice(10, int(k), replace=False).astype(np.int32) for k in rng.integers(1, 8, 40)]
    _say("batch of 40 plain queries n=120")
    br = G.solve_batch(qs)
    for i, q in enumerate(qs):
        ref = O.dpbf(g, grp, q)
        if ref["status"] == O.INFEASIBLE:
            assert br.status[i] == STATUS_INFEASIBLE
            continue
        assert br.costs[i] == ref["cost"], (i, br.costs[i], ref["cost"])
        _check_tree(g, grp, q, br.costs[i], br.tree(i))
    Ds = rng.integers(1, 5, len(qs)).astype(np.int32)
    _say("batch of 40 hop queries n=120")
    bh = G.solve_batch(qs, D=Ds)
    for i, q in enumerate(qs):
        ref = O.hop(g, grp, q, int(Ds[i]))
        if ref["status"] == O.INFEASIBLE:
            assert bh.status[i] == STATUS_INFEASIBLE
            continue
        assert bh.costs[i] == ref["cost"], (i, bh.costs[i], ref["cost"])
        _check_tree(g, grp, q, bh.costs[i], bh.tree(i), int(Ds[i]))


def test_calls_of_different_shapes_on_one_graph(gpu):
    # Batches and single queries of different k and D reuse the same device
    # buffers with different slot layouts: stamps and labels left behind by
    # one call must never change the next call's answer.
    g, grp = workloads.small(150, 400, 10, seed_=17, whi=60, member_p=0.04)
    G = _graph(g, grp)
    rng = np.random.default_rng(21)
    qs = [rng.choice(10, int(k), replace=False).astype(np.int32) for k in rng.integers(2, 8, 24)]
    refs = [O.dpbf(g, grp, q)["cost"] for q in qs]
    for rnd in range(2):
        _say(f"mixed calls round {rnd}: batch of 24, then singles, then hop singles")
        br = G.solve_batch(qs)
        for i in range(len(qs)):
            if refs[i] < O.INF():
                assert br.costs[i] == refs[i], ("batch", rnd, i)
        for i in reversed(range(len(qs))):
            r = G.solve(qs[i])
            if refs[i] < O.INF():
                assert r.cost == refs[i], ("single", rnd, i, r.cost, refs[i])
        for i in range(0, len(qs), 6):
            D = 1 + i % 4
            r = G.solve(qs[i], D=D)
            ref = O.hop(g, grp, qs[i], D)
            assert r.status == (STATUS_OK if ref["status"] == O.OK else STATUS_INFEASIBLE)
            if ref["status"] == O.OK:
                assert r.cost == ref["cost"], ("hop", rnd, i, D)


@pytest.mark.parametrize("div", [0, 4, 64, -1])
def test_dp_band_schedules(gpu, div):
    # A* bands with near/far lists (div > 0), unordered (0), automatic (-1):
    # the schedule changes the work, never the cost
    g, grp = workloads.small(200, 600, 10, seed_=31, whi=80, member_p=0.03)
    G = _graph(g, grp)
    G.set_option("dp_delta_div", div)
    rng = np.random.default_rng(37)
    qs = [rng.choice(10, int(k), replace=False).astype(np.int32) for k in rng.integers(2, 8, 48)]
    _say(f"band schedule div={div}: batch of 48 plain queries n={g.n}")
    br = G.solve_batch(qs)
    for i, q in enumerate(qs):
        ref = O.dpbf(g, grp, q)
        if ref["status"] == O.INFEASIBLE:
            assert br.status[i] == STATUS_INFEASIBLE
            continue
        assert br.costs[i] == ref["cost"], (div, i, br.costs[i], ref["cost"])
        _check_tree(g, grp, q, br.costs[i], br.tree(i))


@pytest.mark.parametrize("div", [0, 4, 64])
def test_hop_band_schedules(gpu, div):
    # hop-bounded queries under A* bands (the first band sized from the
    # smallest key, no upper bound being known) and unordered
    g, grp = workloads.small(150, 420, 8, seed_=47, whi=60, member_p=0.04)
    G = _graph(g, grp)
    G.set_option("dp_delta_div", div)
    rng = np.random.default_rng(53)
    qs = [rng.choice(8, int(k), replace=False).astype(np.int32) for k in rng.integers(2, 7, 40)]
    D = rng.integers(1, 5, 40).astype(np.int32)
    _say(f"hop band schedule div={div}: batch of 40 hop queries n={g.n}")
    br = G.solve_batch(qs, D=D)
    for i, q in enumerate(qs):
        ref = O.hop(g, grp, q, int(D[i]))
        if ref["status"] == O.INFEASIBLE:
            assert br.status[i] == STATUS_INFEASIBLE
            continue
        assert br.costs[i] == ref["cost"], (div, i, br.costs[i], ref["cost"])
        _check_tree(g, grp, q, br.costs[i], br.tree(i), int(D[i]))


@pytest.mark.parametrize("band", [0, 1, 16])
def test_hop_one_label_bands(gpu, band):
    # the one-label phase of hop-bounded queries (no radius) unordered and in
    # doubling bands: labels must end exact either way, or the completion
    # bounds would prune optimal trees
    g, grp = workloads.small(160, 450, 8, seed_=59, whi=90, member_p=0.04)
    G = _graph(g, grp)
    G.set_option("sssp_delta", band)
    rng = np.random.default_rng(61)
    qs = [rng.choice(8, int(k), replace=False).astype(np.int32) for k in rng.integers(2, 8, 36)]
    D = rng.integers(1, 6, 36).astype(np.int32)
    _say(f"hop one-label bands={band}: batch of 36 hop queries n={g.n}")
    br = G.solve_batch(qs, D=D)
    for i, q in enumerate(qs):
        ref = O.hop(g, grp, q, int(D[i]))
        if ref["status"] == O.INFEASIBLE:
            assert br.status[i] == STATUS_INFEASIBLE
            continue
        assert br.costs[i] == ref["cost"], (band, i, br.costs[i], ref["cost"])
        _check_tree(g, grp, q, br.costs[i], br.tree(i), int(D[i]))
    _say(f"hop one-label bands={band}: single query")
    r = G.solve(qs[0], D=int(D[0]))
    ref = O.hop(g, grp, qs[0], int(D[0]))
    if ref["status"] == O.OK:
        assert r.cost == ref["cost"]


@pytest.mark.parametrize("lane,per_item", [(1, 0), (0, 1), (1, 1)])
def test_ablation_switches(gpu, lane, per_item):
    # the reference's ablation variants as switches: per-lane appends and
    # per-item edge walks change the schedule, never the result (a small
    # chunk makes hub rows split into several items)
    g, grp = workloads.small(180, 520, 8, seed_=67, whi=70, member_p=0.04)
    G = _graph(g, grp)
    G.set_option("lane_appends", lane)
    G.set_option("per_item_edges", per_item)
    G.set_option("chunk", 64)
    rng = np.random.default_rng(71)
    qs = [rng.choice(8, int(k), replace=False).astype(np.int32) for k in rng.integers(2, 8, 40)]
    D = rng.integers(1, 5, 40).astype(np.int32)
    _say(f"ablation lane={lane} per_item={per_item}: 40 plain + 40 hop queries n={g.n}")
    bp = G.solve_batch(qs)
    bh = G.solve_batch(qs, D=D)
    for i, q in enumerate(qs):
        for res, ref, Dq in ((bp, O.dpbf(g, grp, q), -1), (bh, O.hop(g, grp, q, int(D[i])), int(D[i]))):
            if ref["status"] == O.INFEASIBLE:
                assert res.status[i] == STATUS_INFEASIBLE
                continue
            assert res.costs[i] == ref["cost"], (lane, per_item, i, Dq, res.costs[i], ref["cost"])
            _check_tree(g, grp, q, res.costs[i], res.tree(i), Dq)


@pytest.mark.parametrize("pct", [0, 40, 85, 100, 150])
def test_one_label_radius_never_changes_a_cost(gpu, pct):
    # the radius only trades one-label work against bound strength (0 is the
    # automatic rule: 85% for up to six groups, 95% for seven); the grow cut
    # by each subset's smallest bound is exercised at every radius
    g, grp = workloads.small(220, 700, 8, seed_=73, whi=90, member_p=0.04)
    G = _graph(g, grp)
    G.set_option("radius_pct", pct)
    rng = np.random.default_rng(79)
    qs = [rng.choice(8, int(k), replace=False).astype(np.int32) for k in rng.integers(2, 8, 40)]
    _say(f"one-label radius {pct}%: 40 plain queries n={g.n}")
    b = G.solve_batch(qs)
    for i, q in enumerate(qs):
        ref = O.dpbf(g, grp, q)
        if ref["status"] == O.INFEASIBLE:
            assert b.status[i] == STATUS_INFEASIBLE
            continue
        assert b.costs[i] == ref["cost"], (pct, i, b.costs[i], ref["cost"])
        _check_tree(g, grp, q, b.costs[i], b.tree(i), -1)
    r = G.solve(qs[0])
    assert r.cost == b.costs[0]


def test_list_overflow_reruns_exactly(gpu):
    # lists far too small for the batch: some list (near, far, one-label or DP)
    # overflows, the batch is re-run with full-size lists, results stay exact;
    # the budget admits the batch at once but not its full-size re-run, which
    # goes in sub-batches
    g, grp = workloads.small(300, 900, 10, seed_=41, whi=50, member_p=0.03)
    G = _graph(g, grp)
    G.set_option("list_permille", 1)
    rng = np.random.default_rng(43)
    qs = [rng.choice(10, int(k), replace=False).astype(np.int32) for k in rng.integers(3, 8, 40)]
    budget = sum(G.footprint(len(q)) for q in qs)
    _say(f"overflowing lists: batch of 40 plain queries n={g.n} budget={budget}")
    br = G.solve_batch(qs, hbm_budget=budget)
    assert br.stats["reruns"] >= 1
    assert br.stats["batches"] >= 3
    for i, q in enumerate(qs):
        ref = O.dpbf(g, grp, q)
        if ref["status"] == O.INFEASIBLE:
            assert br.status[i] == STATUS_INFEASIBLE
            continue
        assert br.costs[i] == ref["cost"], (i, br.costs[i], ref["cost"])
        _check_tree(g, grp, q, br.costs[i], br.tree(i))
    _say("overflowing lists: single query")
    r = G.solve(qs[0])
    assert r.cost == O.dpbf(g, grp, qs[0])["cost"]


def test_infeasible_and_trivial(gpu):
    # two components: {0,1,2} and {3,4}
    g = gen.from_edges(5, [0, 1, 3], [1, 2, 4], [4, 5, 6])
    grp = gen.groups_from_lists([[0], [2], [4], [1, 2], [0, 1, 2]])
    G = _graph(g, grp)
    _say("infeasible / trivial cases")
    r = G.solve(np.array([0, 2], np.int32))
    assert r.status == STATUS_INFEASIBLE and len(r.edges) == 0
    r = G.solve(np.array([0, 1], np.int32))
    assert r.status == STATUS_OK and r.cost == 9
    r = G.solve(np.array([3], np.int32))
    assert r.cost == 0 and len(r.edges) == 0
    r = G.solve(np.array([1, 3, 4], np.int32))  # vertex 2 is in groups 1, 3 and 4
    assert r.cost == 0
    r = G.solve(np.array([0, 1], np.int32), D=1)
    assert r.status == STATUS_INFEASIBLE
    r = G.solve(np.array([0, 1], np.int32), D=2)
    assert r.cost == 9
    G.set_option("infeasible_cost", 100000)
    assert G.solve(np.array([0, 2], np.int32)).cost == 100000


def test_edge_cases_and_refused_queries(gpu):
    # an empty group (infeasible, plain and hop), the largest hop bound,
    # k = 7 hop-bounded, a batch mixing them, and queries the solver refuses
    from gpu4gst import GSTError
    g, grp = workloads.small(40, 100, 8, seed_=83, whi=30, member_p=0.15)
    lists = [grp.members(i).tolist() for i in range(grp.count)] + [[]]
    grp = gen.groups_from_lists(lists)
    empty = grp.count - 1
    full = [i for i in range(empty) if len(lists[i])]
    G = _graph(g, grp)
    cases = [(np.array([full[0], empty], np.int32), -1),
             (np.array([empty, full[1]], np.int32), 2),
             (np.array(full[:6], np.int32), 15),
             (np.array(full[:7], np.int32), 3),
             (np.array(full[:7], np.int32), -1)]
    for q, D in cases:
        _say(f"edge case q={q.tolist()} D={D}")
        r = G.solve(q, D=D)
        ref = O.dpbf(g, grp, q) if D < 0 else O.hop(g, grp, q, D)
        if ref["status"] == O.INFEASIBLE:
            assert r.status == STATUS_INFEASIBLE
            continue
        assert r.cost == ref["cost"], (q.tolist(), D, r.cost, ref["cost"])
        _check_tree(g, grp, q, r.cost, r.edges, D)
    _say("edge cases as one hop batch")
    hq = [q for q, D in cases if D >= 0]
    hD = np.array([D for q, D in cases if D >= 0], np.int32)
    b = G.solve_batch(hq, D=hD)
    for i, q in enumerate(hq):
        ref = O.hop(g, grp, q, int(hD[i]))
        assert (b.status[i] == STATUS_OK) == (ref["status"] == O.OK)
        if ref["status"] == O.OK:
            assert b.costs[i] == ref["cost"]
    for q, D in ((np.array([full[0], full[1], full[0]], np.int32), -1),
                 (np.arange(8, dtype=np.int32), -1),
                 (np.array(full[:2], np.int32), 16),
                 (np.array([grp.count], np.int32), -1)):
        _say(f"refused q={q.tolist()} D={D}")
        with pytest.raises(GSTError):
            G.solve(q, D=D)


def test_full_size_workload_properties(gpu):
    # The 168k-vertex, 6.8M-edge Chung-Lu workload at full size, through
    # properties that do not need the oracle's full solve: every tree is a
    # valid group Steiner tree weighing exactly its cost, batched and single
    # costs agree, a cost is never below the distance between two of its
    # groups (the tree holds a path between them), and two k = 3 queries
    # equal restated PrunedDP++ bit for bit.
    w = workloads.config(2)
    G = _graph(w.graph, w.groups)
    qs = w.queries[:96]
    _say(f"full-size workload {w.name} n={w.graph.n} m={w.graph.m}: batch of 96")
    b = G.solve_batch(qs)
    for i, q in enumerate(qs):
        assert b.status[i] == STATUS_OK
        _check_tree(w.graph, w.groups, q, b.costs[i], b.tree(i))
    for i in range(0, 96, 24):
        _say(f"full-size single query {i} k={len(qs[i])}")
        assert G.solve(qs[i]).cost == b.costs[i]
    for i in range(0, 96, 32):
        q = qs[i]
        dist = [O.sssp(w.graph, w.groups, int(gid))[0] for gid in q]
        pair = max(int(dist[a][w.groups.members(int(q[c]))].min())
                   for a in range(len(q)) for c in range(len(q)) if a != c)
        assert b.costs[i] >= pair, (i, b.costs[i], pair)
    k3 = [i for i, q in enumerate(qs) if len(q) == 3][:2]
    for i in k3:
        _say(f"full-size query {i} vs PrunedDP++")
        assert b.costs[i] == O.prunedpp(w.graph, w.groups, qs[i])["cost"]


def test_full_size_hop_workload_properties(gpu):
    # The 978k-vertex hop-bounded workload at full size: trees valid within
    # their hop bound and weighing their cost, batched = single, and costs
    # monotone in the bound: plain optimum <= cost(D = 5) <= cost(4) <= cost(3).
    w = workloads.config(3)
    G = _graph(w.graph, w.groups)
    qs = w.queries[:48]
    D = np.asarray(w.D[:48], np.int32)
    _say(f"full-size hop workload {w.name} n={w.graph.n} m={w.graph.m}: batch of 48")
    b = G.solve_batch(qs, D=D)
    for i, q in enumerate(qs):
        if b.status[i] == STATUS_OK:
            _check_tree(w.graph, w.groups, q, b.costs[i], b.tree(i), int(D[i]))
    for i in range(0, 48, 16):
        _say(f"full-size hop query {i}: single, D = 3, 4, 5 and plain")
        assert G.solve(qs[i], D=int(D[i])).cost == b.costs[i]
        c = [G.solve(qs[i], D=d) for d in (3, 4, 5)]
        plain = G.solve(qs[i])
        feas = [r.cost for r in c if r.status == STATUS_OK]
        assert feas == sorted(feas, reverse=True)
        assert all(plain.cost <= x for x in feas)


def test_rmat15_queries_vs_prunedpp(gpu):
    w = workloads.config(1, n_queries=24)
    G = _graph(w.graph, w.groups)
    _say(f"rmat15 n={w.graph.n} m={w.graph.m} 24 queries k=3 (batch)")
    br = G.solve_batch(w.queries)
    for i, q in enumerate(w.queries):
        ref = O.prunedpp(w.graph, w.groups, q)
        assert br.costs[i] == ref["cost"], (i, br.costs[i], ref["cost"])
        _check_tree(w.graph, w.groups, q, br.costs[i], br.tree(i))
    for i in range(4):
        _say(f"rmat15 single query {i}")
        r = G.solve(w.queries[i])
        assert r.cost == br.costs[i]


@pytest.mark.parametrize("seed", range(4))
def test_hub_rows_split_into_chunks(gpu, seed):
    """Rows longer than one chunk are queued as several items (virtual split)."""
    rng = np.random.default_rng(900 + seed)
    g = gen.chung_lu(400, 2000, 2.1, 200.0, seed=seed, wlo=1, whi=int(rng.choice([5, 50])))
    grp = gen.groups_zipf(g.n, 12, 1.2, 1, 30, seed=seed)
    G = _graph(g, grp)
    G.set_option("chunk", 64)
    qs = [rng.choice(12, int(k), replace=False).astype(np.int32) for k in rng.integers(2, 8, 8)]
    _say(f"hub graph seed={seed} n={g.n} maxdeg={int(np.diff(g.beg).max())} chunk=64")
    br = G.solve_batch(qs)
    Ds = rng.integers(1, 5, len(qs)).astype(np.int32)
    bh = G.solve_batch(qs, D=Ds)
    for i, q in enumerate(qs):
        ref = O.dpbf(g, grp, q)
        assert br.costs[i] == ref["cost"]
        _check_tree(g, grp, q, br.costs[i], br.tree(i))
        rh = O.hop(g, grp, q, int(Ds[i]))
        if rh["status"] == O.INFEASIBLE:
            assert bh.status[i] == STATUS_INFEASIBLE
        else:
            assert bh.costs[i] == rh["cost"]
            _check_tree(g, grp, q, bh.costs[i], bh.tree(i), int(Ds[i]))
