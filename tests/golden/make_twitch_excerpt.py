"""Writes the Twitch fixtures under tests/golden/twitch/ from the reference's
data directory (data/Twitch_beg_pos.bin, Twitch.g, Twitch{3,5,7}.csv; the
edge and weight files are not shipped with the reference).

* Twitch_beg_pos.bin and Twitch{3,5,7}.csv are copied unchanged.
* Twitch_excerpt.g keeps, unchanged, the lines of Twitch.g for every group
  that the first QUERIES queries of each query file name (other lines are
  left out; readers treat missing ids as empty groups).
* twitch_facts.json records counts read from the full files (lines, groups,
  largest member id) that the excerpt alone cannot show.

Run from the repository root: python tests/golden/make_twitch_excerpt.py
"""
import json
import os
import shutil

SRC = "/root/reference/data"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "twitch")
QUERIES = 10


def main():
    os.makedirs(DST, exist_ok=True)
    for name in ("Twitch_beg_pos.bin", "Twitch3.csv", "Twitch5.csv", "Twitch7.csv"):
        shutil.copyfile(os.path.join(SRC, name), os.path.join(DST, name))
    need = set()
    facts = {"query_lines": {}}
    for k in (3, 5, 7):
        with open(os.path.join(SRC, f"Twitch{k}.csv")) as f:
            lines = [ln.split() for ln in f if ln.strip()]
        facts["query_lines"][str(k)] = len(lines)
        for t in lines[:QUERIES]:
            need.update(int(x) for x in t)
    kept, n_lines, max_member, max_id = [], 0, -1, 0
    with open(os.path.join(SRC, "Twitch.g")) as f:
        for line in f:
            if not line.strip():
                continue
            n_lines += 1
            head, _, rest = line.partition(":")
            gid = int(head[1:]) - 1
            max_id = max(max_id, gid)
            mem = [int(x) for x in rest.split()]
            if mem:
                max_member = max(max_member, max(mem))
            if gid in need:
                kept.append(line if line.endswith("\n") else line + "\n")
    with open(os.path.join(DST, "Twitch_excerpt.g"), "w") as f:
        f.writelines(kept)
    facts.update({"group_lines": n_lines, "max_group_id": max_id, "max_member": max_member,
                  "excerpt_groups": sorted(need), "queries_per_file_in_excerpt": QUERIES})
    with open(os.path.join(DST, "twitch_facts.json"), "w") as f:
        json.dump(facts, f, indent=1)


if __name__ == "__main__":
    main()
