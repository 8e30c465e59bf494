"""gst_run (gfx950 build) on the reference's Twitch inputs: the real group
file (excerpt) and query files, with a configuration-model graph on the real
Twitch degree sequence (Twitch_beg_pos.bin) since the reference does not ship
Twitch's edges. Costs in the result CSVs must equal the oracle's.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from gpu4gst import formats, workloads
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TW = os.path.join(ROOT, "tests", "golden", "twitch")
EXE = os.path.join(ROOT, "gpu4gst-sigmod-fullexp_amd", "gpu4gst", "gst_run")
SEED = 0x67535400 + 16 * 6


@pytest.fixture(scope="module")
def twitch_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("twitch")
    deg = np.diff(np.fromfile(os.path.join(TW, "Twitch_beg_pos.bin"), "<i8"))
    g = workloads.degree_sequence_graph(deg, seed_=SEED, wlo=1, whi=100)
    formats.write_bin(str(d / "Twitch"), g.beg, g.col, g.cw)
    formats.write_in(str(d / "Twitch.in"), g.n, g.src, g.dst, g.w)
    shutil.copyfile(os.path.join(TW, "Twitch_excerpt.g"), d / "Twitch.g")
    for k in (3, 5, 7):
        shutil.copyfile(os.path.join(TW, f"Twitch{k}.csv"), d / f"Twitch{k}.csv")
    return str(d) + "/", g, formats.read_groups(os.path.join(TW, "Twitch_excerpt.g"))


def _run(args):
    print("gst_run", " ".join(args), flush=True)
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("k", [3, 5, 7])
def test_twitch_wb_plain(gpu, twitch_dir, k):
    d, g, grp = twitch_dir
    _run(["TrimCDP-WB", "2", d, "Twitch", str(k), "0", "9"])
    rows = formats.read_result_csv(formats.result_file(d, "TrimCDP-WB", "Twitch", k, 0, 9))
    qs = formats.read_queries(d + f"Twitch{k}.csv")
    assert len(rows) == 10
    for row in rows:
        q = qs[int(row["task_ID"])]
        assert row["task"] == " ".join(map(str, q.tolist())) + " "
        assert int(row["GPU2_nonHop_cost"]) == O.prunedpp(g, grp, q)["cost"]


def test_twitch_hop_and_text_input(gpu, twitch_dir):
    d, g, grp = twitch_dir
    _run(["D-TrimCDP-WB", "2", d, "Twitch", "3", "3", "0", "9"])
    rows = formats.read_result_csv(formats.result_file(d, "D-TrimCDP-WB", "Twitch", 3, 0, 9, 3))
    qs = formats.read_queries(d + "Twitch3.csv")
    for row in rows:
        ref = O.hop(g, grp, qs[int(row["task_ID"])], 3)["cost"]
        assert int(row["GPU2_Hop_cost"]) == (ref if ref < O.INF() else 100000)
    # the text graph and one query per call give the same costs
    _run(["TrimCDP", "2", d, "Twitch", "5", "0", "4", "--single"])
    rows = formats.read_result_csv(formats.result_file(d, "TrimCDP", "Twitch", 5, 0, 4))
    qs = formats.read_queries(d + "Twitch5.csv")
    for row in rows:
        assert int(row["GPU_nonHop_cost"]) == O.prunedpp(g, grp, qs[int(row["task_ID"])])["cost"]
