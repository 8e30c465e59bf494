"""Python mirror of the solver call surface over libgst.so.

``GSTGraph`` owns one device graph (CSR + groups). ``solve`` is the drop-in of
the reference's per-query solvers (DP_GPU, code/TrimCDP/src/DPQ.cu:292, and
DP_gpu, code/D-TrimCDP/src/DPQ.cu:147) with the tree output the CPU solvers
have (graph_v_of_v_idealID_PrunedDPPlusPlus, code/PrunedDP++/include/
CPUNONHOP.h:496; HOP_cpu, code/D-PrunedDP++/include/CPUHOP.h:107);
``solve_batch`` replaces the reference's one-query-at-a-time harness loops
(code/TrimCDP/include/exp_GPU_nonHop.h:40-144).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib

# per-query status (include/gst.h GST_Q_*): optimal tree found; no tree covers
# the groups; the optimum (if any) lies beyond the int32 cost range; cost exact
# but the tree too long for the recovery buffer
STATUS_OK, STATUS_INFEASIBLE, STATUS_RANGE, STATUS_NO_TREE = 0, 1, 2, 3


class GSTError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gst error {code}: {msg}")
        self.code = code


def _check(rc: int):
    if rc != _lib.OK:
        raise GSTError(rc, _lib.lib().gst_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclass
class SolveResult:
    cost: int
    status: int
    edges: np.ndarray  # int32 [n, 2] vertex pairs
    stats: dict = field(default_factory=dict)


@dataclass
class BatchResult:
    costs: np.ndarray     # int32 [nq]
    status: np.ndarray    # int32 [nq]
    offsets: np.ndarray   # int64 [nq+1]
    edges: np.ndarray     # int32 [total, 2]
    edge_ids: np.ndarray  # int32 [total] undirected edge ids
    stats: dict = field(default_factory=dict)

    def tree(self, q: int) -> np.ndarray:
        return self.edges[self.offsets[q]:self.offsets[q + 1]]

    def tree_ids(self, q: int) -> np.ndarray:
        return self.edge_ids[self.offsets[q]:self.offsets[q + 1]]


class GSTGraph:
    """Device-resident graph. ``beg``/``col``/``w`` form a symmetric CSR
    (int64 or int32 arrays, the reference's *_beg_pos / *_csr / *_weight.bin
    layout); ``eid`` optionally maps each CSR entry to an undirected edge id."""

    def __init__(self, beg, col, w, eid=None, num_undirected=None):
        L = _lib.lib()
        beg = np.ascontiguousarray(beg)
        bits = 64 if beg.dtype == np.int64 else 32
        dt = np.int64 if bits == 64 else np.int32
        beg = np.ascontiguousarray(beg, dt)
        col = np.ascontiguousarray(col, dt)
        w = np.ascontiguousarray(w, dt)
        self.V = int(beg.shape[0] - 1)
        self.E = int(col.shape[0])
        h = C.c_void_p()
        _check(L.gst_graph_create(_p(beg), _p(col), _p(w), self.V, self.E, bits, C.byref(h)))
        self._h = h
        self.eid = None
        self.num_undirected = None
        if eid is not None:
            eid = np.ascontiguousarray(eid, np.int32)
            n_und = int(num_undirected if num_undirected is not None else eid.max() + 1)
            _check(L.gst_graph_set_edge_ids(self._h, _p(eid), n_und))
            self.eid = eid
            self.num_undirected = n_und
        self.G = 0

    @classmethod
    def from_generated(cls, g) -> "GSTGraph":
        """From a gpu4gst.gen.Graph (CSR with undirected edge ids)."""
        return cls(g.beg, g.col, g.cw, g.eid, g.m)

    def set_groups(self, off, mem=None):
        """Groups as CSR (offsets int64 [G+1], members int32) or a gen.Groups."""
        if mem is None:
            off, mem = off.off, off.mem
        off = np.ascontiguousarray(off, np.int64)
        mem = np.ascontiguousarray(mem, np.int32)
        _check(_lib.lib().gst_groups_set(self._h, _p(off), _p(mem), off.shape[0] - 1))
        self.G = int(off.shape[0] - 1)

    def set_weights_csr(self, w):
        w = np.ascontiguousarray(w, np.int32)
        if w.shape[0] != self.E:
            raise ValueError("weights must have one entry per CSR entry")
        _check(_lib.lib().gst_graph_set_weights(self._h, _p(w)))

    def update_weights(self, csr_index, w):
        """weight[csr_index[i]] = w[i]; both entries of an edge must change together."""
        idx = np.ascontiguousarray(csr_index, np.int64)
        w = np.ascontiguousarray(np.broadcast_to(np.asarray(w, np.int32), idx.shape), np.int32)
        _check(_lib.lib().gst_graph_update_weights(self._h, _p(idx), _p(w), idx.shape[0]))

    def set_option(self, name: str, value: int):
        _check(_lib.lib().gst_set_option(self._h, _lib.OPT[name], int(value)))

    def footprint(self, k: int, D: int = -1) -> int:
        return int(_lib.lib().gst_query_footprint(self._h, k, D))

    def solve(self, query, D: int = -1, edge_cap: int = 1 << 16) -> SolveResult:
        q = np.ascontiguousarray(query, np.int32)
        cost, ne, status = C.c_int32(0), C.c_int32(0), C.c_int32(0)
        uv = np.zeros(2 * edge_cap, np.int32)
        st = _lib.GstStats()
        _check(_lib.lib().gst_solve(self._h, _p(q), q.shape[0], D, C.byref(cost), _p(uv), edge_cap,
                                    C.byref(ne), C.byref(status), C.byref(st)))
        return SolveResult(cost.value, status.value, uv[:2 * ne.value].reshape(-1, 2).copy(),
                           st.as_dict())

    def solve_batch(self, queries, D=None, hbm_budget: int = 0, edge_cap: int | None = None
                    ) -> BatchResult:
        qs = [np.ascontiguousarray(q, np.int32) for q in queries]
        nq = len(qs)
        off = np.zeros(nq + 1, np.int64)
        for i, q in enumerate(qs):
            off[i + 1] = off[i] + q.shape[0]
        groups = np.concatenate(qs) if nq else np.zeros(0, np.int32)
        Dv = None
        if D is not None:
            Dv = np.ascontiguousarray(np.broadcast_to(np.asarray(D, np.int32), (nq,)), np.int32)
        costs = np.zeros(nq, np.int32)
        status = np.zeros(nq, np.int32)
        eoff = np.zeros(nq + 1, np.int64)
        cap = edge_cap if edge_cap is not None else 64 * max(nq, 1)
        L = _lib.lib()
        while True:
            uv = np.zeros(2 * cap, np.int32)
            ids = np.zeros(cap, np.int32)
            st = _lib.GstStats()
            rc = L.gst_solve_batch(self._h, _p(off), _p(groups), _p(Dv), nq, hbm_budget,
                                   _p(costs), _p(status), _p(eoff), _p(uv), _p(ids), cap,
                                   C.byref(st))
            if rc == _lib.ERR_CAPACITY and edge_cap is None:
                cap = int(eoff[-1]) + 1
                continue
            _check(rc)
            break
        tot = int(eoff[-1])
        return BatchResult(costs, status, eoff, uv[:2 * tot].reshape(-1, 2).copy(),
                           ids[:tot].copy(), st.as_dict())

    def autotune(self, queries, D=None, batch_size: int = 0, max_seconds: float = 10.0) -> dict:
        """Times candidate kernel settings on `queries` and keeps the fastest for
        their size class on this graph (gst_autotune); batch_size = 1 tunes the
        one-query path. max_seconds <= 0 only reports the settings in force."""
        qs = [np.ascontiguousarray(q, np.int32) for q in queries]
        nq = len(qs)
        off = np.zeros(nq + 1, np.int64)
        for i, q in enumerate(qs):
            off[i + 1] = off[i] + q.shape[0]
        groups = np.concatenate(qs) if nq else np.zeros(0, np.int32)
        Dv = None
        if D is not None:
            Dv = np.ascontiguousarray(np.broadcast_to(np.asarray(D, np.int32), (nq,)), np.int32)
        out = _lib.GstTuning()
        _check(_lib.lib().gst_autotune(self._h, _p(off), _p(groups), _p(Dv), nq, batch_size,
                                       float(max_seconds), C.byref(out)))
        return out.as_dict()

    def trace(self) -> list[dict]:
        """Per-launch trace of the last solve_batch run with options profile=1 and
        trace=1: phase, batch, items, active slots, slots, ms."""
        L = _lib.lib()
        n = C.c_int64(0)
        rc = L.gst_trace_get(self._h, None, 0, C.byref(n))  # the length only
        if rc != _lib.ERR_CAPACITY:
            _check(rc)
        rows = (_lib.GstTraceRow * max(1, n.value))()
        _check(L.gst_trace_get(self._h, rows, n.value, C.byref(n)))
        return [{f: getattr(r, f) for f, _ in _lib.GstTraceRow._fields_ if f != "pad"}
                for r in rows[:n.value]]

    def backward(self, nq: int, d_offsets: int, d_ids: int, d_grad_out: int, d_grad_w: int,
                 num_undirected: int):
        """Device pointers (ints); see gst_backward."""
        _check(_lib.lib().gst_backward(self._h, nq, C.c_void_p(d_offsets), C.c_void_p(d_ids),
                                       C.c_void_p(d_grad_out), C.c_void_p(d_grad_w),
                                       num_undirected))

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().gst_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dp_gpu(graph: GSTGraph, query_groups) -> SolveResult:
    """Plain TrimCDP for one query (the role of DP_GPU, code/TrimCDP/src/DPQ.cu:292)."""
    return graph.solve(query_groups, -1)


def dp_gpu_hop(graph: GSTGraph, query_groups, D: int) -> SolveResult:
    """Hop-bounded D-TrimCDP for one query (DP_gpu, code/D-TrimCDP/src/DPQ.cu:147)."""
    return graph.solve(query_groups, D)
