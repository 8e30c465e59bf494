"""The allocation-free entry point gst_solve_noalloc, through the C ABI.

Its per-query buffers live in a static device array that every call on the
GPU shares and that keeps the last call's contents; each call must clear what
has to start empty. Checked against gst_solve (the allocating entry point) on
two calls in a row, the second with a smaller input (a smaller graph, fewer
groups), for plain and hop-bounded queries, and a query too large for the
array is refused with GST_ERR_OOM.
"""
import ctypes as C

import numpy as np
import pytest

from gpu4gst import GSTGraph, workloads, _lib
import oracle as O

pytestmark = pytest.mark.gpu


def _noalloc(G, q, D=-1, cap=4096):
    L = _lib.lib()
    q = np.ascontiguousarray(q, np.int32)
    cost, ne, status = C.c_int32(0), C.c_int32(0), C.c_int32(0)
    uv = np.zeros(2 * cap, np.int32)
    rc = L.gst_solve_noalloc(G._h, q.ctypes.data_as(C.c_void_p), q.shape[0], D, C.byref(cost),
                             uv.ctypes.data_as(C.c_void_p), cap, C.byref(ne), C.byref(status),
                             None)
    return rc, cost.value, status.value, uv[:2 * ne.value].reshape(-1, 2)


def _check(g, grp, G, q, D):
    print(f"noalloc n={g.n} m={g.m} q={list(q)} D={D}", flush=True)
    rc, cost, status, edges = _noalloc(G, q, D)
    assert rc == _lib.OK, _lib.lib().gst_last_error().decode()
    ref = G.solve(q, D=D)
    assert (cost, status) == (ref.cost, ref.status)
    if status == 0 and len(edges):
        code, wsum = O.verify_tree(g, grp, q, edges, D)
        assert code == 0 and wsum == cost
    return cost


def test_noalloc_matches_solve_twice_smaller_second(gpu):
    big, bgrp = workloads.small(400, 1200, 8, seed_=11, whi=50, member_p=0.02)
    small, sgrp = workloads.small(30, 60, 4, seed_=12, whi=9, member_p=0.1)
    Gb = GSTGraph.from_generated(big)
    Gb.set_groups(bgrp)
    Gs = GSTGraph.from_generated(small)
    Gs.set_groups(sgrp)
    for D in (-1, 3):
        c1 = _check(big, bgrp, Gb, np.array([0, 2, 4, 5, 7], np.int32), D)
        c2 = _check(small, sgrp, Gs, np.array([1, 3], np.int32), D)
        assert c1 == O.dpbf(big, bgrp, np.array([0, 2, 4, 5, 7], np.int32))["cost"] or D >= 0
        assert c2 == (O.dpbf(small, sgrp, np.array([1, 3], np.int32))["cost"] if D < 0 else
                      O.hop(small, sgrp, np.array([1, 3], np.int32), D)["cost"])
    # the allocating path still works after the allocation-free calls
    r = Gb.solve_batch([np.array([0, 1, 2], np.int32), np.array([3, 4], np.int32)])
    for i, q in enumerate(([0, 1, 2], [3, 4])):
        assert int(r.costs[i]) == O.dpbf(big, bgrp, np.array(q, np.int32))["cost"]
    Gb.close()
    Gs.close()


def test_noalloc_refuses_queries_beyond_the_array(gpu):
    w = workloads.config(2)
    G = GSTGraph.from_generated(w.graph)
    G.set_groups(w.groups)
    L = _lib.lib()
    need = L.gst_noalloc_bytes(G._h, 7, -1)
    print(f"k=7 on V={w.graph.n} needs {need} bytes of scratch", flush=True)
    assert need > (256 << 20)
    rc, *_ = _noalloc(G, w.queries[2])  # k = 7
    assert rc == _lib.ERR_OOM
    assert b"scratch" in L.gst_last_error()
    assert L.gst_noalloc_bytes(G._h, 3, -1) <= (256 << 20)
    q = w.queries[0]  # k = 3 fits
    rc, cost, status, _ = _noalloc(G, q)
    assert rc == _lib.OK and cost == G.solve(q).cost
    G.close()
