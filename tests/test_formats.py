"""Reference on-disk formats: the Python readers (gpu4gst.formats) and the C++
readers of gst_run (include/gst_formats.hpp) on the reference's own Twitch
files (tests/golden/twitch, written by tests/golden/make_twitch_excerpt.py)
and on generated graphs; gst_run's four command lines, built against the CPU
emulation of the kernels, against the oracle.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from gpu4gst import formats, gen, workloads
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TW = os.path.join(ROOT, "tests", "golden", "twitch")
INC = os.path.join(ROOT, "include")
EMU = os.path.join(ROOT, "gpu4gst-sigmod-fullexp_amd", "_build", "libgst_emu.so")


@pytest.fixture(scope="module")
def dump(tmp_path_factory):
    exe = tmp_path_factory.mktemp("fd") / "formats_dump"
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-I", INC,
                        os.path.join(ROOT, "tests", "cpp", "formats_dump.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

    def run(mode, path):
        out = subprocess.run([str(exe), mode, path], capture_output=True, text=True)
        assert out.returncode == 0, out.stderr
        return out.stdout.split()
    return run


def _graph_digest(beg, col, w):
    col = np.asarray(col, np.int64)
    w = np.asarray(w, np.int64)
    i = np.arange(col.shape[0])
    return ["graph", str(beg.shape[0] - 1), str(int(beg[-1])), str(int(col.sum())),
            str(int(w.sum())), str(int(((i % 7 + 1) * col).sum()))]


def test_twitch_files(dump):
    facts = json.load(open(os.path.join(TW, "twitch_facts.json")))
    beg = np.fromfile(os.path.join(TW, "Twitch_beg_pos.bin"), "<i8")
    deg = np.diff(beg)
    assert beg.shape[0] - 1 == 34118 and beg[0] == 0
    assert int(beg[-1]) == 858226 and int(beg[-1]) % 2 == 0  # both directions stored
    assert deg.min() >= 1 and facts["max_member"] < beg.shape[0] - 1
    for k in (3, 5, 7):
        q = formats.read_queries(os.path.join(TW, f"Twitch{k}.csv"))
        assert len(q) == facts["query_lines"][str(k)] == 300
        assert all(len(x) == k and len(set(x.tolist())) == k for x in q)
        assert max(int(x.max()) for x in q) <= facts["max_group_id"]
        d = dump("q", os.path.join(TW, f"Twitch{k}.csv"))
        ids = np.concatenate(q).astype(np.int64)
        sx = sum(int(((np.arange(len(x)) + 1) * x.astype(np.int64)).sum()) for x in q)
        assert d == ["queries", str(len(q)), str(ids.shape[0]), str(int(ids.sum())), str(sx)]
    grp = formats.read_groups(os.path.join(TW, "Twitch_excerpt.g"))
    nonempty = [i for i in range(grp.count) if grp.members(i).shape[0]]
    assert nonempty == facts["excerpt_groups"]
    assert grp.count == max(nonempty) + 1
    for i in nonempty:
        m = grp.members(i)
        assert (np.diff(m) > 0).all() and m.max() <= facts["max_member"]
    d = dump("g", os.path.join(TW, "Twitch_excerpt.g"))
    sx = sum(int(((np.arange(len(grp.members(i))) % 5 + 1) * grp.members(i).astype(np.int64)).sum())
             for i in range(grp.count))
    assert d == ["groups", str(grp.count), str(len(nonempty)), str(grp.mem.shape[0]),
                 str(int(grp.mem.astype(np.int64).sum())), str(sx)]


def test_bin_round_trip_and_edge_pairing(tmp_path, dump):
    g = gen.rmat(10, 3000, 0.57, 0.19, 0.19, seed=11, wlo=1, whi=50)
    prefix = str(tmp_path / "r10")
    formats.write_bin(prefix, g.beg, g.col, g.cw)
    beg, col, w = formats.read_bin(prefix)
    assert np.array_equal(beg, g.beg) and np.array_equal(col, g.col) and np.array_equal(w, g.cw)
    assert dump("bin", prefix) == _graph_digest(beg, col, w) + ["1"]
    eid, m = formats.pair_edge_ids(beg, col, w)
    assert m == g.m
    # pairing is a relabelling of the generator's own undirected ids
    mp = {}
    for a, b in zip(eid.tolist(), g.eid.tolist()):
        assert mp.setdefault(a, b) == b
    assert len(set(mp.values())) == m
    # a CSR whose two directions disagree is refused
    w2 = w.copy()
    w2[0] += 1
    with pytest.raises(ValueError):
        formats.pair_edge_ids(beg, col, w2)


def test_in_round_trip(tmp_path, dump):
    g = workloads.degree_sequence_graph(
        np.diff(np.fromfile(os.path.join(TW, "Twitch_beg_pos.bin"), "<i8"))[:3000] % 40 + 1,
        seed_=5, wlo=1, whi=100)
    p = str(tmp_path / "x.in")
    # a parallel edge and a weight tie exercise the (neighbour, weight) row order
    src = np.concatenate([g.src, [g.src[0]]])
    dst = np.concatenate([g.dst, [g.dst[0]]])
    w = np.concatenate([g.w, [1]])
    formats.write_in(p, g.n, src, dst, w)
    n, s2, d2, w2 = formats.read_in(p)
    assert n == g.n and np.array_equal(s2, src) and np.array_equal(d2, dst)
    beg, col, cw, eid = formats.csr_from_edges(n, s2, d2, w2)
    d = dump("in", p)
    assert d == _graph_digest(beg, col, cw) + ["1"]
    # rows ordered by (neighbour, weight)
    for v in range(0, n, 97):
        r = list(zip(col[beg[v]:beg[v + 1]].tolist(), cw[beg[v]:beg[v + 1]].tolist()))
        assert r == sorted(r)
    # without the extra edge the CSR is the generator's (its rows are neighbour-sorted)
    beg0, col0, cw0, _ = formats.csr_from_edges(g.n, g.src, g.dst, g.w)
    assert np.array_equal(beg0, g.beg) and np.array_equal(col0, g.col) and np.array_equal(
        cw0, g.cw)


def test_groups_queries_round_trip(tmp_path):
    grp = gen.groups_zipf(500, 40, 1.2, 1, 60, seed=9)
    p = str(tmp_path / "x.g")
    formats.write_groups(p, grp)
    back = formats.read_groups(p)
    assert np.array_equal(back.off, grp.off)
    assert all(np.array_equal(back.members(i), np.sort(grp.members(i))) for i in range(40))
    qs = gen.queries(grp, np.array([3, 5, 7, 1], np.int32), seed=4)
    formats.write_queries(str(tmp_path / "x5.csv"), qs)
    assert all(np.array_equal(a, b) for a, b in zip(formats.read_queries(str(tmp_path / "x5.csv")),
                                                     qs))
    (tmp_path / "bad.g").write_text("group1: 1 2\n")
    with pytest.raises(ValueError):
        formats.read_groups(str(tmp_path / "bad.g"))


def _cli(tmp_path, lib):
    exe = tmp_path / "gst_run"
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-I", INC,
                        os.path.join(ROOT, "gpu4gst-sigmod-fullexp_amd", "csrc", "gst_cli.cpp"),
                        lib, f"-Wl,-rpath,{os.path.dirname(lib)}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


def test_cli_variants_on_emulated_kernels(tmp_path):
    if not os.path.exists(EMU):
        pytest.skip("libgst_emu.so not built (run __graft_entry__.build())")
    exe = _cli(tmp_path, EMU)
    g, grp = workloads.small(26, 45, 6, seed_=31, whi=9, member_p=0.15)
    d = str(tmp_path) + "/"
    formats.write_in(d + "s.in", g.n, g.src, g.dst, g.w)
    formats.write_bin(d + "s", g.beg, g.col.astype(np.int64), g.cw.astype(np.int64))
    formats.write_groups(d + "s.g", grp)
    rng = np.random.default_rng(3)
    qs = [rng.choice(6, 3, replace=False) for _ in range(4)]
    formats.write_queries(d + "s3.csv", qs)
    for variant, D in (("TrimCDP", None), ("TrimCDP-WB", None), ("D-TrimCDP", 2),
                       ("D-TrimCDP-WB", 1)):
        args = [exe, variant, "2", d, "s", "3"] + ([str(D)] if D is not None else []) + ["1", "3"]
        r = subprocess.run(args, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        rows = formats.read_result_csv(formats.result_file(d, variant, "s", 3, 1, 3, D))
        assert [int(x["task_ID"]) for x in rows] == [1, 2, 3]
        cost_col = [c for c in rows[0] if c.endswith("_cost")][0]
        for row in rows:
            q = qs[int(row["task_ID"])]
            assert row["task"] == " ".join(map(str, q.tolist())) + " "
            ref = O.dpbf(g, grp, q) if D is None else O.hop(g, grp, q, D)
            want = ref["cost"]
            if want >= O.INF():
                want = 0 if variant == "TrimCDP" else 100000
            assert int(row[cost_col]) == want, (variant, row)
    # --single and the symlink form
    link = tmp_path / "TrimCDP"
    os.symlink(exe, link)
    r = subprocess.run([str(link), "2", d, "s", "3", "0", "0", "--single"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = formats.read_result_csv(formats.result_file(d, "TrimCDP", "s", 3, 0, 0))
    assert int(rows[0]["GPU_nonHop_cost"]) == O.dpbf(g, grp, qs[0])["cost"]
    # a range past the end of the query file is an error, not a crash
    r = subprocess.run([exe, "TrimCDP", "2", d, "s", "3", "0", "9"], capture_output=True, text=True)
    assert r.returncode == 1 and "outside" in r.stderr
