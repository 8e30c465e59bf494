"""ctypes bindings of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker, never by the product
path. The restated reference functions and their citations are listed in the
header of gst_oracle.cpp.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OK, INFEASIBLE, INVALID, CAPACITY = 0, 1, 2, 3


def build():
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle.so"], check=True)


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp, i32 = C.c_void_p, C.c_int32
        L.orc_inf.restype = i32
        common = [i32, vp, vp, vp, i32, vp, vp, i32, vp]
        L.orc_sssp.argtypes = [i32, vp, vp, vp, vp, vp, i32, vp, vp]
        L.orc_prunedpp.argtypes = common + [vp, vp, i32, vp, vp, vp, vp]
        L.orc_dpbf.argtypes = common + [vp, vp, i32, vp]
        L.orc_hop.argtypes = common + [i32, vp, vp, i32, vp, vp, vp]
        L.orc_trimcdp_mt.argtypes = common + [i32, vp, vp, vp, vp]
        L.orc_hop_mt.argtypes = common + [i32, i32, vp, vp, vp]
        L.orc_brute_plain.argtypes = common + [vp]
        L.orc_brute_hop.argtypes = common + [i32, vp]
        L.orc_verify_tree.argtypes = common + [i32, i32, vp, i32, vp]
        _LIB = L
    return _LIB


def INF() -> int:
    return int(_lib().orc_inf())


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(g, groups, q):
    q = np.ascontiguousarray(q, np.int32)
    return [g.n, _p(g.beg), _p(g.col), _p(g.cw), groups.count, _p(groups.off), _p(groups.mem),
            q.shape[0], _p(q)], q


def sssp(g, groups, gid: int):
    dist = np.empty(g.n, np.int32)
    pred = np.empty(g.n, np.int32)
    _lib().orc_sssp(g.n, _p(g.beg), _p(g.col), _p(g.cw), _p(groups.off), _p(groups.mem), gid,
                    _p(dist), _p(pred))
    return dist, pred


def _edges_out(cap):
    return np.zeros(2 * cap, np.int32), C.c_int32(0), C.c_int32(0)


def prunedpp(g, groups, q, cap=1 << 16):
    """Restated PrunedDP++; returns dict(status, cost, edges, dp_s, total_s, popped, finite)."""
    args, q = _c(g, groups, q)
    e, ne, cost = _edges_out(cap)
    dp, tot = C.c_double(0), C.c_double(0)
    st = np.zeros(2, np.int64)
    rc = _lib().orc_prunedpp(*args, C.byref(cost), _p(e), cap, C.byref(ne), C.byref(dp),
                             C.byref(tot), _p(st))
    return dict(status=rc, cost=cost.value, edges=e[:2 * ne.value].reshape(-1, 2).copy(),
                dp_s=dp.value, total_s=tot.value, popped=int(st[0]), finite=int(st[1]))


def dpbf(g, groups, q, cap=1 << 16):
    args, q = _c(g, groups, q)
    e, ne, cost = _edges_out(cap)
    rc = _lib().orc_dpbf(*args, C.byref(cost), _p(e), cap, C.byref(ne))
    return dict(status=rc, cost=cost.value, edges=e[:2 * ne.value].reshape(-1, 2).copy())


def hop(g, groups, q, D: int, cap=1 << 16):
    args, q = _c(g, groups, q)
    e, ne, cost = _edges_out(cap)
    dp = C.c_double(0)
    st = np.zeros(2, np.int64)
    rc = _lib().orc_hop(*args, D, C.byref(cost), _p(e), cap, C.byref(ne), C.byref(dp), _p(st))
    return dict(status=rc, cost=cost.value, edges=e[:2 * ne.value].reshape(-1, 2).copy(),
                dp_s=dp.value, popped=int(st[0]), finite=int(st[1]))


def trimcdp_mt(g, groups, q, threads: int = 0):
    args, q = _c(g, groups, q)
    cost = C.c_int32(0)
    dp, tot = C.c_double(0), C.c_double(0)
    st = np.zeros(2, np.int64)
    rc = _lib().orc_trimcdp_mt(*args, threads, C.byref(cost), C.byref(dp), C.byref(tot), _p(st))
    return dict(status=rc, cost=cost.value, dp_s=dp.value, total_s=tot.value,
                processed=int(st[0]), levels=int(st[1]))


def hop_mt(g, groups, q, D: int, threads: int = 0):
    args, q = _c(g, groups, q)
    cost = C.c_int32(0)
    dp = C.c_double(0)
    st = np.zeros(2, np.int64)
    rc = _lib().orc_hop_mt(*args, D, threads, C.byref(cost), C.byref(dp), _p(st))
    return dict(status=rc, cost=cost.value, dp_s=dp.value, processed=int(st[0]),
                levels=int(st[1]))


def brute_plain(g, groups, q):
    args, q = _c(g, groups, q)
    cost = C.c_int32(0)
    rc = _lib().orc_brute_plain(*args, C.byref(cost))
    return rc, cost.value


def brute_hop(g, groups, q, D: int):
    args, q = _c(g, groups, q)
    cost = C.c_int32(0)
    rc = _lib().orc_brute_hop(*args, D, C.byref(cost))
    return rc, cost.value


def verify_tree(g, groups, q, edges, D: int = -1, root_hint: int = -1):
    """0 = feasible; see orc_verify_tree for the codes. Returns (code, weight)."""
    args, q = _c(g, groups, q)
    edges = np.ascontiguousarray(np.asarray(edges, np.int32).reshape(-1, 2))
    wsum = C.c_int64(0)
    rc = _lib().orc_verify_tree(*args, D, edges.shape[0], _p(edges), root_hint, C.byref(wsum))
    return rc, wsum.value
