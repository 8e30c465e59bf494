"""Seeded synthetic graphs, groups and queries (ctypes over libgstgen.so).

The reference has no generator on its live path and its only dataset ships
without edge files (data/.MISSING_LARGE_BLOBS), so every workload is built
here from a seed. Graphs are undirected and simple; ``csr()`` returns the
symmetric CSR the solver takes (rows sorted by neighbour id, as the
reference loader sorts them at
code/TrimCDP/include/graph_hash_of_mixed_weighted_read_for_GSTP.h:130-133).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libgstgen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        vp, i32, i64, u64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
        L.gg_rmat.restype = vp
        L.gg_rmat.argtypes = [i32, i64, dbl, dbl, dbl, u64]
        L.gg_chung_lu.restype = vp
        L.gg_chung_lu.argtypes = [i64, i64, dbl, dbl, u64]
        L.gg_from_edges.restype = vp
        L.gg_from_edges.argtypes = [i64, i64, vp, vp, vp]
        L.gg_keep_lcc.argtypes = [vp]
        L.gg_lcc_vertices.restype = i64
        L.gg_lcc_vertices.argtypes = [vp, vp]
        L.gg_set_weights_uniform.argtypes = [vp, i32, i32, u64]
        L.gg_num_vertices.restype = i64
        L.gg_num_vertices.argtypes = [vp]
        L.gg_num_edges.restype = i64
        L.gg_num_edges.argtypes = [vp]
        L.gg_get_edges.argtypes = [vp, vp, vp, vp]
        L.gg_get_csr.argtypes = [vp, vp, vp, vp, vp]
        L.gg_free.argtypes = [vp]
        L.gg_groups_poisson.restype = vp
        L.gg_groups_poisson.argtypes = [i64, i32, dbl, u64]
        L.gg_groups_zipf.restype = vp
        L.gg_groups_zipf.argtypes = [i64, i32, dbl, i32, i32, u64]
        L.gg_groups_zipf_pool.restype = vp
        L.gg_groups_zipf_pool.argtypes = [i64, vp, i64, i32, dbl, i32, i32, u64]
        L.gg_groups_count.restype = i64
        L.gg_groups_count.argtypes = [vp]
        L.gg_groups_total.restype = i64
        L.gg_groups_total.argtypes = [vp]
        L.gg_groups_get.argtypes = [vp, vp, vp]
        L.gg_groups_free.argtypes = [vp]
        L.gg_queries.restype = i32
        L.gg_queries.argtypes = [i32, vp, i32, vp, u64, vp]
        L.gg_uniform_ints.argtypes = [i64, i32, i32, u64, vp]
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class Graph:
    """Undirected simple graph: edge list (src < dst) plus symmetric CSR."""

    n: int
    src: np.ndarray  # int32 [E]
    dst: np.ndarray  # int32 [E]
    w: np.ndarray  # int32 [E]
    beg: np.ndarray  # int64 [V+1]
    col: np.ndarray  # int32 [2E]
    cw: np.ndarray  # int32 [2E]
    eid: np.ndarray  # int32 [2E] undirected edge id of each CSR entry
    # vertices of the largest component (increasing ids); set when the graph
    # keeps all its vertices (lcc=False), None when it is the component itself
    lcc: np.ndarray | None = None

    @property
    def m(self) -> int:
        return int(self.src.shape[0])


def _finish(h, lo: int, hi: int, wseed: int, lcc: bool, pool: bool = False) -> Graph:
    L = _lib()
    comp = None
    try:
        if lcc:
            L.gg_keep_lcc(h)
        elif pool:
            nc = L.gg_lcc_vertices(h, None)
            comp = np.empty(nc, np.int32)
            L.gg_lcc_vertices(h, _p(comp))
        if lo is not None:
            L.gg_set_weights_uniform(h, lo, hi, wseed)
        n = L.gg_num_vertices(h)
        m = L.gg_num_edges(h)
        src = np.empty(m, np.int32)
        dst = np.empty(m, np.int32)
        w = np.empty(m, np.int32)
        L.gg_get_edges(h, _p(src), _p(dst), _p(w))
        beg = np.empty(n + 1, np.int64)
        col = np.empty(2 * m, np.int32)
        cw = np.empty(2 * m, np.int32)
        eid = np.empty(2 * m, np.int32)
        L.gg_get_csr(h, _p(beg), _p(col), _p(cw), _p(eid))
    finally:
        L.gg_free(h)
    return Graph(int(n), src, dst, w, beg, col, cw, eid, comp)


def rmat(scale: int, m: int, a=0.57, b=0.19, c=0.19, *, seed: int, wlo=1, whi=100,
         wseed: int | None = None, lcc=True) -> Graph:
    """lcc=True keeps the largest component (relabelled 0..V'-1); lcc=False
    keeps all 2^scale vertices and records the component in Graph.lcc."""
    h = _lib().gg_rmat(scale, m, a, b, c, seed)
    if not h:
        raise ValueError("invalid R-MAT parameters")
    return _finish(h, wlo, whi, seed + 1 if wseed is None else wseed, lcc, pool=not lcc)


def chung_lu(n: int, m: int, beta: float, dmax: float, *, seed: int, wlo=1, whi=1000,
             wseed: int | None = None, lcc=True) -> Graph:
    """lcc as for rmat()."""
    h = _lib().gg_chung_lu(n, m, beta, dmax, seed)
    if not h:
        raise ValueError("invalid Chung-Lu parameters")
    return _finish(h, wlo, whi, seed + 1 if wseed is None else wseed, lcc, pool=not lcc)


def from_edges(n: int, src, dst, w=None, lcc=False) -> Graph:
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    wp = None
    if w is not None:
        w = np.ascontiguousarray(w, np.int32)
        wp = _p(w)
    h = _lib().gg_from_edges(n, src.shape[0], _p(src), _p(dst), wp)
    if not h:
        raise ValueError("invalid edge list")
    return _finish(h, None, None, 0, lcc)


@dataclass
class Groups:
    off: np.ndarray  # int64 [G+1]
    mem: np.ndarray  # int32 [total]

    @property
    def count(self) -> int:
        return int(self.off.shape[0] - 1)

    def members(self, g: int) -> np.ndarray:
        return self.mem[self.off[g]:self.off[g + 1]]


def _groups(h) -> Groups:
    L = _lib()
    try:
        G = L.gg_groups_count(h)
        tot = L.gg_groups_total(h)
        off = np.empty(G + 1, np.int64)
        mem = np.empty(tot, np.int32)
        L.gg_groups_get(h, _p(off), _p(mem))
    finally:
        L.gg_groups_free(h)
    return Groups(off, mem)


def groups_poisson(n: int, G: int, lam: float, *, seed: int) -> Groups:
    h = _lib().gg_groups_poisson(n, G, lam, seed)
    if not h:
        raise ValueError("invalid group parameters")
    return _groups(h)


def groups_zipf(n: int, G: int, s: float, smin: int, smax: int, *, seed: int) -> Groups:
    h = _lib().gg_groups_zipf(n, G, s, smin, smax, seed)
    if not h:
        raise ValueError("invalid group parameters")
    return _groups(h)


def groups_zipf_pool(n: int, pool, G: int, s: float, smin: int, smax: int, *,
                     seed: int) -> Groups:
    """groups_zipf with members drawn from the vertex ids in `pool` only."""
    pool = np.ascontiguousarray(pool, np.int32)
    h = _lib().gg_groups_zipf_pool(n, _p(pool), pool.shape[0], G, s, smin, smax, seed)
    if not h:
        raise ValueError("invalid group parameters")
    return _groups(h)


def groups_from_lists(lists) -> Groups:
    off = np.zeros(len(lists) + 1, np.int64)
    for i, l in enumerate(lists):
        off[i + 1] = off[i] + len(l)
    mem = np.array([v for l in lists for v in sorted(l)], np.int32)
    return Groups(off, mem)


def queries(groups: Groups, ks, *, seed: int) -> list[np.ndarray]:
    ks = np.ascontiguousarray(ks, np.int32)
    out = np.empty(int(ks.sum()), np.int32)
    rc = _lib().gg_queries(groups.count, _p(groups.off), ks.shape[0], _p(ks), seed, _p(out))
    if rc != 0:
        raise ValueError("not enough non-empty groups for the requested k")
    res, pos = [], 0
    for k in ks:
        res.append(out[pos:pos + k].copy())
        pos += k
    return res


def uniform_ints(count: int, lo: int, hi: int, *, seed: int) -> np.ndarray:
    out = np.empty(count, np.int32)
    _lib().gg_uniform_ints(count, lo, hi, seed, _p(out))
    return out
