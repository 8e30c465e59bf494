"""The CPU oracle pinned against brute force, closed forms and invariants.

orc_dpbf / orc_prunedpp / orc_trimcdp_mt (plain) and orc_hop / orc_hop_mt
(hop-bounded) must equal exhaustive search on tiny graphs; k = 1 and k = 2 reduce to
textbook answers; the hop optimum is non-increasing in D and equals the plain
optimum once D >= V - 1. The tree verifier must reject planted mistakes.
"""
import numpy as np
import pytest

from gpu4gst import gen, workloads
import oracle as O

INF = None


def _tiny(rng, nmax=12, emax=20):
    n = int(rng.integers(2, nmax + 1))
    m = int(rng.integers(0, emax))
    perm = rng.permutation(n)
    src = np.concatenate([perm[:-1], rng.integers(0, n, m)])
    dst = np.concatenate([perm[1:], rng.integers(0, n, m)])
    w = rng.integers(1, int(rng.choice([2, 5, 30])), src.shape[0])
    g = gen.from_edges(n, src, dst, w)
    G = int(rng.integers(1, 6))
    lists = [sorted(set(rng.integers(0, n, int(rng.integers(1, 4))).tolist())) for _ in range(G)]
    return g, gen.groups_from_lists(lists)


def test_plain_matches_brute_force():
    rng = np.random.default_rng(7)
    checked = 0
    for _ in range(160):
        g, grp = _tiny(rng)
        k = int(rng.integers(1, grp.count + 1))
        q = rng.choice(grp.count, k, replace=False).astype(np.int32)
        rc, cb = O.brute_plain(g, grp, q)
        assert rc == O.OK
        for r in (O.dpbf(g, grp, q), O.prunedpp(g, grp, q)):
            assert r["cost"] == cb
            if len(r["edges"]):
                code, wsum = O.verify_tree(g, grp, q, r["edges"])
                assert code == 0 and wsum == cb
        assert O.trimcdp_mt(g, grp, q, 2)["cost"] == cb
        checked += 1
    assert checked == 160


def test_hop_matches_brute_force():
    rng = np.random.default_rng(8)
    for _ in range(120):
        g, grp = _tiny(rng, nmax=9, emax=12)
        if g.m > 22:
            continue
        k = int(rng.integers(1, grp.count + 1))
        q = rng.choice(grp.count, k, replace=False).astype(np.int32)
        for D in range(0, 5):
            rb, cb = O.brute_hop(g, grp, q, D)
            r = O.hop(g, grp, q, D)
            assert r["status"] == rb
            assert O.hop_mt(g, grp, q, D, 2)["status"] == rb
            if rb == O.OK:
                assert O.hop_mt(g, grp, q, D, 2)["cost"] == cb
                assert r["cost"] == cb
                if len(r["edges"]):
                    code, wsum = O.verify_tree(g, grp, q, r["edges"], D)
                    assert code == 0 and wsum == cb


def test_closed_forms_k1_k2():
    rng = np.random.default_rng(9)
    for seed in range(12):
        g, grp = workloads.small(40, 60, 5, seed_=seed, whi=25, member_p=0.1)
        for a in range(5):
            assert O.dpbf(g, grp, [a])["cost"] == 0
            dist, _ = O.sssp(g, grp, a)
            for b in range(5):
                if a == b:
                    continue
                want = int(dist[grp.members(b)].min())  # distance between the two groups
                assert O.dpbf(g, grp, [a, b])["cost"] == want
                assert O.prunedpp(g, grp, [a, b])["cost"] == want


def test_vertex_in_every_group_costs_zero():
    g = gen.from_edges(4, [0, 1, 2], [1, 2, 3], [5, 5, 5])
    grp = gen.groups_from_lists([[0, 2], [2, 3], [1, 2]])
    for f in (O.dpbf, O.prunedpp):
        r = f(g, grp, [0, 1, 2])
        assert r["cost"] == 0 and len(r["edges"]) == 0
    assert O.hop(g, grp, [0, 1, 2], 0)["cost"] == 0


def test_path_and_star_closed_forms():
    # path 0-1-2-3-4 with weights 1,2,3,4; groups at the ends -> whole path
    g = gen.from_edges(5, [0, 1, 2, 3], [1, 2, 3, 4], [1, 2, 3, 4])
    grp = gen.groups_from_lists([[0], [4], [2]])
    assert O.dpbf(g, grp, [0, 1])["cost"] == 10
    assert O.hop(g, grp, [0, 1], 4)["cost"] == 10
    assert O.hop(g, grp, [0, 1], 3)["status"] == O.INFEASIBLE
    # star: centre 0, leaves 1..4 with weights 3,4,5,6, one group per leaf
    g = gen.from_edges(5, [0, 0, 0, 0], [1, 2, 3, 4], [3, 4, 5, 6])
    grp = gen.groups_from_lists([[1], [2], [3], [4]])
    assert O.dpbf(g, grp, [0, 1, 2, 3])["cost"] == 18
    assert O.prunedpp(g, grp, [0, 1, 2, 3])["cost"] == 18
    assert O.hop(g, grp, [0, 1, 2, 3], 2)["cost"] == 18
    assert O.hop(g, grp, [0, 1, 2, 3], 1)["status"] == O.INFEASIBLE


def test_hop_monotone_in_D_and_reaches_plain():
    for seed in range(8):
        g, grp = workloads.small(14, 10, 4, seed_=50 + seed, whi=9, member_p=0.15)
        q = [0, 1, 2, 3]
        plain = O.dpbf(g, grp, q)["cost"]
        prev = None
        for D in range(0, g.n):
            r = O.hop(g, grp, q, D)
            c = r["cost"] if r["status"] == O.OK else None
            if prev is not None:
                assert c is not None and c <= prev
            prev = c if c is not None else prev
        assert prev == plain


def test_medium_graphs_all_oracles_agree():
    for seed in range(10):
        g, grp = workloads.small(300, 700, 8, seed_=seed, whi=100, member_p=0.02)
        rng = np.random.default_rng(seed)
        for k in (3, 5, 6):
            q = rng.choice(8, k, replace=False).astype(np.int32)
            a = O.dpbf(g, grp, q)["cost"]
            assert O.prunedpp(g, grp, q)["cost"] == a
            assert O.trimcdp_mt(g, grp, q, 4)["cost"] == a
            for D in (2, 4):
                h = O.hop(g, grp, q, D)
                assert O.hop_mt(g, grp, q, D, 4)["cost"] == h["cost"]


def test_verifier_rejects_planted_mistakes():
    g = gen.from_edges(6, [0, 1, 2, 3, 0], [1, 2, 3, 4, 5], [1, 1, 1, 1, 1])
    grp = gen.groups_from_lists([[0], [4], [5]])
    q = [0, 1, 2]
    good = np.array([[0, 1], [1, 2], [2, 3], [3, 4], [0, 5]])
    assert O.verify_tree(g, grp, q, good) == (0, 5)
    assert O.verify_tree(g, grp, q, good[:-1])[0] == 4            # group 2 not touched
    assert O.verify_tree(g, grp, q, good[[0, 1, 3, 4]])[0] == 3   # disconnected
    assert O.verify_tree(g, grp, q, np.vstack([good, [[1, 2]]]))[0] == 2  # duplicate edge
    assert O.verify_tree(g, grp, q, np.vstack([good[:-1], [[0, 4]]]))[0] == 1  # not an edge
    assert O.verify_tree(g, grp, q, good, D=3)[0] == 5            # diameter 5 > 3
    assert O.verify_tree(g, grp, q, good, D=5)[0] == 0
