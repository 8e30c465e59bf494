"""Rolling admission on the GPU: a call whose tables do not fit the HBM budget
at once runs the one-label phase for every query and then refills DP table
regions as queries converge. Results must equal the lockstep batches'
(GST_OPT_ROLLING = 0) bit for bit, and the oracle's.

Each case prints its name and shape before the GPU call.
"""
import numpy as np
import pytest

from gpu4gst import GSTGraph, gen, workloads, STATUS_INFEASIBLE, STATUS_OK
import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _verify(g, grp, q, cost, edges, D=-1):
    if len(edges) == 0:
        gs = np.zeros(g.n, np.int64)
        for i, gid in enumerate(q):
            gs[grp.members(int(gid))] |= 1 << i
        assert cost == 0 and (gs == (1 << len(q)) - 1).any()
        return
    code, wsum = O.verify_tree(g, grp, q, edges, D)
    assert code == 0 and wsum == cost, (code, wsum, cost)


def _powerlaw(n, m, seed):
    g = gen.chung_lu(n, m, 2.2, n / 8.0, seed=seed, wlo=1, whi=200, wseed=seed + 1)
    grp = gen.groups_zipf(g.n, 400, 1.2, 1, max(2, g.n // 50), seed=seed + 2)
    return g, grp


@pytest.mark.parametrize("hop", [False, True])
def test_rolling_equals_lockstep_and_oracle(gpu, hop):
    g, grp = _powerlaw(6000, 30000, 71 if hop else 73)
    G = GSTGraph.from_generated(g)
    G.set_groups(grp)
    rng = np.random.default_rng(75)
    ks = [3, 4, 5] if hop else [3, 5, 7]
    qs = gen.queries(grp, np.array([ks[i % 3] for i in range(48)], np.int32), seed=77)
    D = gen.uniform_ints(len(qs), 2, 4, seed=79) if hop else None
    foot = sum(G.footprint(len(q), -1 if D is None else int(D[i])) for i, q in enumerate(qs))
    budget = foot // 5
    _say(f"rolling hop={hop} n={g.n} m={g.m} queries={len(qs)} budget={budget}")
    roll = G.solve_batch(qs, D=D, hbm_budget=budget)
    st = roll.stats
    _say(f"rolling: {st['rolling_queries']} queries, {st['batches']} admission rounds, "
         f"at most {st['max_resident']} resident")
    assert st["rolling_queries"] == len(qs)
    assert 1 <= st["max_resident"] < len(qs)
    assert st["batches"] >= 2
    _say("lockstep batches of the same call")
    G.set_option("rolling", 0)
    lock = G.solve_batch(qs, D=D, hbm_budget=budget)
    G.set_option("rolling", 1)
    assert lock.stats["rolling_queries"] == 0 and lock.stats["batches"] >= 2
    assert np.array_equal(roll.costs, lock.costs)
    assert np.array_equal(roll.status, lock.status)
    for i, q in enumerate(qs):
        Dq = -1 if D is None else int(D[i])
        if roll.status[i] == STATUS_OK:
            _verify(g, grp, q, int(roll.costs[i]), roll.tree(i), Dq)
    for i in rng.choice(len(qs), 6, replace=False):
        q = qs[i]
        Dq = -1 if D is None else int(D[i])
        _say(f"oracle query {i} k={len(q)} D={Dq}")
        ref = O.hop(g, grp, q, Dq) if hop else O.prunedpp(g, grp, q)
        if ref["status"] == O.INFEASIBLE:
            assert roll.status[i] == STATUS_INFEASIBLE
        else:
            assert int(roll.costs[i]) == ref["cost"], (i, int(roll.costs[i]), ref["cost"])
    G.close()


@pytest.mark.parametrize("div", [-1, 0, 16])
def test_rolling_refills_under_a_resident_cap(gpu, div):
    # GST_OPT_MAX_BATCH caps the queries holding tables at once: with 3 of 30
    # resident every query after the third enters a region another one left
    g, grp = workloads.small(400, 1600, 10, seed_=81, whi=60, member_p=0.05)
    G = GSTGraph.from_generated(g)
    G.set_groups(grp)
    rng = np.random.default_rng(83)
    qs = [rng.choice(10, int(k), replace=False).astype(np.int32) for k in rng.integers(2, 8, 30)]
    G.set_option("dp_delta_div", div)
    G.set_option("max_batch", 3)
    _say(f"resident cap 3, dp_delta_div={div}, {len(qs)} plain queries n={g.n}")
    r = G.solve_batch(qs)
    assert r.stats["rolling_queries"] == len(qs) and r.stats["max_resident"] <= 3
    for i, q in enumerate(qs):
        ref = O.dpbf(g, grp, q)
        if ref["status"] == O.INFEASIBLE:
            assert r.status[i] == STATUS_INFEASIBLE
            continue
        assert int(r.costs[i]) == ref["cost"], (i, int(r.costs[i]), ref["cost"])
        _verify(g, grp, q, int(r.costs[i]), r.tree(i))
    D = np.array([int(x) for x in rng.integers(1, 4, len(qs))], np.int32)
    _say(f"resident cap 3, dp_delta_div={div}, {len(qs)} hop queries")
    rh = G.solve_batch(qs, D=D)
    for i, q in enumerate(qs):
        ref = O.hop(g, grp, q, int(D[i]))
        if ref["status"] == O.INFEASIBLE:
            assert rh.status[i] == STATUS_INFEASIBLE
            continue
        assert int(rh.costs[i]) == ref["cost"], (i, int(rh.costs[i]), ref["cost"])
        _verify(g, grp, q, int(rh.costs[i]), rh.tree(i), int(D[i]))
    G.close()


def test_rolling_overflow_reruns_the_rest(gpu):
    # worklists far too small: the rolling run overflows, the queries it had
    # not collected yet are solved again with full-size lists
    g, grp = workloads.small(300, 900, 10, seed_=85, whi=50, member_p=0.03)
    G = GSTGraph.from_generated(g)
    G.set_groups(grp)
    G.set_option("list_permille", 1)
    G.set_option("max_batch", 4)
    rng = np.random.default_rng(87)
    qs = [rng.choice(10, int(k), replace=False).astype(np.int32) for k in rng.integers(3, 8, 24)]
    _say(f"rolling with overflowing lists: {len(qs)} plain queries n={g.n}")
    r = G.solve_batch(qs)
    assert r.stats["reruns"] >= 1
    for i, q in enumerate(qs):
        ref = O.dpbf(g, grp, q)
        if ref["status"] == O.INFEASIBLE:
            assert r.status[i] == STATUS_INFEASIBLE
            continue
        assert int(r.costs[i]) == ref["cost"], (i, int(r.costs[i]), ref["cost"])
        _verify(g, grp, q, int(r.costs[i]), r.tree(i))
    G.close()
