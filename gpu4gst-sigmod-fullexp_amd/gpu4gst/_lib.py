"""ctypes declarations of libgst.so (include/gst.h)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# GPU4GST_LIB selects another build of the same library (the CPU emulation
# build used for fault finding); the default is the gfx950 build.
LIB_PATH = os.environ.get("GPU4GST_LIB") or os.path.join(_HERE, "libgst.so")

# symbols declared by include/gst.h, checked at load time
EXPORTS = ("gst_graph_create", "gst_graph_set_edge_ids", "gst_graph_set_weights",
           "gst_graph_update_weights",
           "gst_groups_set", "gst_set_option", "gst_solve", "gst_solve_batch", "gst_backward",
           "gst_query_footprint", "gst_last_error", "gst_graph_destroy", "gst_version",
           "gst_autotune", "gst_trace_get", "gst_solve_noalloc", "gst_noalloc_bytes")

OK, ERR_INVALID, ERR_OOM, ERR_DEVICE, ERR_CAPACITY, ERR_INTERNAL = 0, 1, 2, 3, 4, 5
OPT = dict(chunk=1, grid=2, iters_per_sync=3, hop_prune=4, profile=5, infeasible_cost=6,
           max_tree_edges=7, use_graphs=8, sssp_delta=10, dp_delta_div=11, radius_pct=12, max_batch=14,
           sssp_slots=15, list_permille=16, lane_appends=17, per_item_edges=18,
           backpointers=19, trace=20, rolling=21)


class GstStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "queries", "batches", "sssp_iterations", "dp_iterations", "sssp_items", "sssp_relax",
        "sssp_updates", "sssp_items_unbounded",
        "states_expanded", "grow_relax", "merge_relax", "improvements", "lb_terms",
        "items_processed", "max_frontier", "bytes_alg")] + [
        ("seconds_total", C.c_double), ("expand_ms", C.c_double), ("expand_launches", C.c_int64),
        ("sssp_expand_ms", C.c_double), ("phase_ms", C.c_double * 3),
        ("sssp_expand_launches", C.c_int64), ("reruns", C.c_int64), ("ub_rows", C.c_int64),
        ("bytes_layout", C.c_int64), ("backpointers_stored", C.c_int64),
        ("bp_fallbacks", C.c_int64), ("rolling_queries", C.c_int64),
        ("max_resident", C.c_int64)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["phase_ms"] = list(self.phase_ms)
        return d


class GstTuning(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("chunk", "grid", "dp_div", "sssp_band", "size_class",
                                         "trials")] + [
        ("seconds_before", C.c_double), ("seconds_after", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class GstTraceRow(C.Structure):
    _fields_ = [("phase", C.c_int32), ("batch", C.c_int32), ("items", C.c_int64),
                ("active_slots", C.c_int32), ("slots", C.c_int32), ("ms", C.c_float),
                ("pad", C.c_int32)]


_LIB = None


def lib():
    """Loads libgst.so; raises if it is missing (there is no fallback path)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        missing = [s for s in EXPORTS if not hasattr(L, s)]
        if missing:
            raise RuntimeError(f"libgst.so lacks symbols {missing}")
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L.gst_version.restype = i32
        if L.gst_version() < 2:  # gst_stats gained fields in version 2
            raise RuntimeError(f"{LIB_PATH} is version {L.gst_version()}; this wrapper needs 2")
        L.gst_last_error.restype = C.c_char_p
        L.gst_graph_create.argtypes = [vp, vp, vp, i64, i64, i32, C.POINTER(vp)]
        L.gst_graph_set_edge_ids.argtypes = [vp, vp, i64]
        L.gst_graph_set_weights.argtypes = [vp, vp]
        L.gst_graph_update_weights.argtypes = [vp, vp, vp, i64]
        L.gst_groups_set.argtypes = [vp, vp, vp, i32]
        L.gst_set_option.argtypes = [vp, i32, i64]
        L.gst_solve.argtypes = [vp, vp, i32, i32, vp, vp, i32, vp, vp, vp]
        L.gst_solve_batch.argtypes = [vp, vp, vp, vp, i32, i64, vp, vp, vp, vp, vp, i64, vp]
        L.gst_backward.argtypes = [vp, i32, vp, vp, vp, vp, i64]
        L.gst_query_footprint.restype = i64
        L.gst_query_footprint.argtypes = [vp, i32, i32]
        L.gst_graph_destroy.argtypes = [vp]
        L.gst_autotune.argtypes = [vp, vp, vp, vp, i32, i32, C.c_double, vp]
        L.gst_trace_get.argtypes = [vp, vp, i64, C.POINTER(i64)]
        L.gst_solve_noalloc.argtypes = [vp, vp, i32, i32, vp, vp, i32, vp, vp, vp]
        L.gst_noalloc_bytes.restype = i64
        L.gst_noalloc_bytes.argtypes = [vp, i32, i32]
        _LIB = L
    return _LIB
