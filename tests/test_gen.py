"""Synthetic workload generators: determinism and structural properties."""
import hashlib

import numpy as np

from gpu4gst import gen, workloads


def _h(*arrays):
    m = hashlib.sha256()
    for a in arrays:
        m.update(np.ascontiguousarray(a).tobytes())
    return m.hexdigest()


def _check_csr(g):
    beg, col, w, eid = g.beg, g.col, g.cw, g.eid
    assert beg[0] == 0 and beg[-1] == 2 * g.m and np.all(np.diff(beg) >= 0)
    for v in range(g.n):
        row = col[beg[v]:beg[v + 1]]
        assert np.all(np.diff(row) > 0)  # sorted, no duplicates
        assert not np.any(row == v)
    # symmetry: (v,u,w) present iff (u,v,w) present
    rows = np.repeat(np.arange(g.n), np.diff(beg))
    fwd = set(zip(rows.tolist(), col.tolist(), w.tolist()))
    bwd = set(zip(col.tolist(), rows.tolist(), w.tolist()))
    assert fwd == bwd
    assert np.array_equal(g.w[eid], w)


def _connected(g):
    seen = np.zeros(g.n, bool)
    stack = [0]
    seen[0] = True
    while stack:
        v = stack.pop()
        for u in g.col[g.beg[v]:g.beg[v + 1]]:
            if not seen[u]:
                seen[u] = True
                stack.append(u)
    return seen.all()


def test_rmat_deterministic_and_connected():
    a = gen.rmat(10, 3000, seed=1)
    b = gen.rmat(10, 3000, seed=1)
    c = gen.rmat(10, 3000, seed=2)
    assert _h(a.src, a.dst, a.w) == _h(b.src, b.dst, b.w)
    assert _h(a.src, a.dst) != _h(c.src, c.dst)
    assert _connected(a)
    assert a.w.min() >= 1 and a.w.max() <= 100
    _check_csr(a)


def test_chung_lu_degrees():
    g = gen.chung_lu(5000, 40000, 2.1, 800.0, seed=3)
    assert _connected(g)
    deg = np.diff(g.beg)
    assert deg.max() > 20 * np.median(deg)  # heavy tail
    assert g.m <= 40000
    _check_csr(g)


def test_groups_and_queries():
    grp = gen.groups_zipf(1000, 200, 1.2, 1, 300, seed=5)
    sz = np.diff(grp.off)
    assert sz.min() >= 1 and sz.max() <= 300
    for i in range(grp.count):
        m = grp.members(i)
        assert np.all(np.diff(m) > 0) and m.min() >= 0 and m.max() < 1000
    p = gen.groups_poisson(1000, 50, 0.5, seed=6)
    assert np.diff(p.off).sum() >= 1000  # every vertex joins >= 1 group
    qs = gen.queries(grp, [3, 5, 7], seed=9)
    assert [len(q) for q in qs] == [3, 5, 7]
    for q in qs:
        assert len(set(q.tolist())) == len(q)
    assert all(np.array_equal(x, y) for x, y in zip(qs, gen.queries(grp, [3, 5, 7], seed=9)))


def test_workload_1_shape():
    w = workloads.config(1)
    assert w.graph.n > 20000 and w.graph.m > 250000
    assert w.groups.count == 2000 and len(w.queries) == 100
    assert all(len(q) == 3 for q in w.queries)
    assert _connected(w.graph)
