"""Readers and writers for the reference's on-disk formats (numpy, host only).

The same formats as include/gst_formats.hpp (the C++ readers of gst_run):

* ``<name>_beg_pos.bin`` / ``_csr.bin`` / ``_weight.bin``: little-endian int64
  beg_pos[V+1], csr[E], weight[E], E = beg_pos[V], both directions of every
  edge (code/TrimCDP-WB/include/graph.hpp:25-113; writer
  data/tuple_text_to_bin.cpp:205-237).
* ``<name>.in``: first line starts with V, then "v1 v2 w" lines; both
  directions are added and every row is sorted by (neighbour, weight)
  (code/TrimCDP/include/graph_hash_of_mixed_weighted_read_for_GSTP.h:75-148).
* ``<name>.g``: "gN:v v ..." lines, group id N-1, members sorted by id
  (same file:190-239; code/TrimCDP-WB/include/read_group.hpp:9-53).
* ``<name><k>.csv``: one query per line, space-separated group ids
  (same file:240-286).
* result CSV of the harnesses (code/TrimCDP/include/exp_GPU_nonHop.h:83-85).
"""
from __future__ import annotations

import csv
import os

import numpy as np

from . import gen


def read_bin(prefix: str):
    """(beg_pos, csr, weight) as int64 arrays; csr/weight cut to beg_pos[V]."""
    beg = np.fromfile(prefix + "_beg_pos.bin", dtype="<i8")
    if beg.shape[0] < 2 or beg[0] != 0 or beg[-1] <= 0 or (np.diff(beg) < 0).any():
        raise ValueError(f"{prefix}_beg_pos.bin: not a CSR offset array")
    E = int(beg[-1])
    col = np.fromfile(prefix + "_csr.bin", dtype="<i8")
    w = np.fromfile(prefix + "_weight.bin", dtype="<i8")
    if col.shape[0] < E or w.shape[0] < E:
        raise ValueError(f"{prefix}: csr/weight files hold fewer than beg_pos[V] entries")
    col, w = col[:E], w[:E]
    if E and (col.min() < 0 or col.max() >= beg.shape[0] - 1):
        raise ValueError(f"{prefix}_csr.bin: vertex id out of range")
    return beg.astype(np.int64), col.astype(np.int64), w.astype(np.int64)


def write_bin(prefix: str, beg, col, w) -> None:
    for suffix, a in (("_beg_pos.bin", beg), ("_csr.bin", col), ("_weight.bin", w)):
        np.ascontiguousarray(a, dtype="<i8").tofile(prefix + suffix)


def csr_from_edges(n: int, src, dst, w):
    """Symmetric CSR the way the reference's text reader builds it: both
    directions of every line, each row sorted by (neighbour, weight). Returns
    (beg int64, col int32, w int32, eid int32) with eid = the line's index."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    w = np.asarray(w, np.int64)
    m = src.shape[0]
    rows = np.concatenate([src, dst])
    cols = np.concatenate([dst, src])
    ws = np.concatenate([w, w])
    ids = np.concatenate([np.arange(m), np.arange(m)])
    order = np.lexsort((ids, ws, cols, rows))
    beg = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=beg[1:])
    return (beg, cols[order].astype(np.int32), ws[order].astype(np.int32),
            ids[order].astype(np.int32))


def read_in(path: str):
    """(n, src, dst, w) of a .in file (every line of exactly three integers)."""
    with open(path) as f:
        head = f.readline().split()
        if not head:
            raise ValueError(f"{path}: first line has no vertex count")
        n = int(float(head[0]))
        src, dst, w = [], [], []
        for lineno, line in enumerate(f, 2):
            t = line.split()
            if len(t) != 3:
                continue
            a, b, c = (int(x) for x in t)
            if not (0 <= a < n and 0 <= b < n):
                raise ValueError(f"{path}: line {lineno} names a vertex out of range")
            src.append(a)
            dst.append(b)
            w.append(c)
    return n, np.array(src, np.int32), np.array(dst, np.int32), np.array(w, np.int32)


def write_in(path: str, n: int, src, dst, w) -> None:
    with open(path, "w") as f:
        f.write(f"{n} {len(src)}\n")
        for a, b, c in zip(np.asarray(src).tolist(), np.asarray(dst).tolist(),
                           np.asarray(w).tolist()):
            f.write(f"{a} {b} {c}\n")


def read_groups(path: str) -> gen.Groups:
    """Groups of a .g file; "gN" is id N-1, missing ids are empty groups."""
    lists: dict[int, list[int]] = {}
    with open(path) as f:
        for lineno, line in enumerate(f, 1):
            line = line.rstrip("\r\n")
            if not line.strip():
                continue
            head, sep, rest = line.partition(":")
            if not sep or not head.startswith("g") or not head[1:].isdigit() or int(head[1:]) < 1:
                raise ValueError(f"{path}: line {lineno} is not gN:members")
            lists[int(head[1:]) - 1] = sorted(int(x) for x in rest.split())
    G = max(lists) + 1 if lists else 0
    return gen.groups_from_lists([lists.get(i, []) for i in range(G)])


def write_groups(path: str, groups: gen.Groups) -> None:
    with open(path, "w") as f:
        for i in range(groups.count):
            f.write(f"g{i + 1}:" + " ".join(map(str, groups.members(i).tolist())) + "\n")


def read_queries(path: str) -> list:
    out = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if t:
                out.append(np.array([int(x) for x in t], np.int32))
    return out


def write_queries(path: str, queries) -> None:
    with open(path, "w") as f:
        for q in queries:
            f.write(" ".join(map(str, np.asarray(q).tolist())) + "\n")


def pair_edge_ids(beg, col, w):
    """Undirected edge ids for a symmetric CSR without them (the .bin files):
    the entry u->v and an entry v->u of the same weight share an id. Returns
    (eid int32 [E], number of undirected edges); raises if the CSR is not
    symmetric."""
    beg = np.asarray(beg, np.int64)
    n = beg.shape[0] - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(beg))
    col = np.asarray(col, np.int64)
    w = np.asarray(w, np.int64)
    lo, hi = np.minimum(rows, col), np.maximum(rows, col)
    fwd = rows <= col  # self-loops: both copies sit in one row, split them below
    loop = rows == col
    idx = np.arange(rows.shape[0])
    a = idx[fwd]
    a = a[np.lexsort((a, w[a], hi[a], lo[a]))]
    b = idx[~fwd]
    b = b[np.lexsort((b, w[b], hi[b], lo[b]))]
    la = a[loop[a]]
    a = a[~loop[a]]
    if la.shape[0] % 2:
        raise ValueError("CSR is not symmetric (odd self-loop copies)")
    a = np.concatenate([a, la[0::2]])
    b = np.concatenate([b, la[1::2]])
    a = a[np.lexsort((a, w[a], hi[a], lo[a]))]
    b = b[np.lexsort((b, w[b], hi[b], lo[b]))]
    if a.shape[0] != b.shape[0] or not (np.array_equal(lo[a], lo[b]) and np.array_equal(
            hi[a], hi[b]) and np.array_equal(w[a], w[b])):
        raise ValueError("CSR is not symmetric")
    eid = np.empty(rows.shape[0], np.int32)
    eid[a] = np.arange(a.shape[0])
    eid[b] = np.arange(b.shape[0])
    return eid, int(a.shape[0])


def read_result_csv(path: str) -> list:
    """Rows of a harness result file as dicts (the trailing empty line skipped)."""
    with open(path, newline="") as f:
        rows = [r for r in csv.DictReader(f) if r and r.get("task_ID")]
    return rows


def result_file(path: str, variant: str, data: str, T: int, first: int, last: int,
                D: int | None = None) -> str:
    name = f"{variant}_{data}_T{T}_" + (f"D{D}_" if variant.startswith("D-") else "")
    return os.path.join(path, "result", name + f"{first}-{last}.csv")
