"""The autotuner on the 168k-vertex / 6.8M-edge workload: the size-class model
picks 256-edge chunks for one-query calls and 4096 for large batches; a
measured tuning pass keeps its choice per size class without changing any
result (costs against the committed oracle costs)."""
import os
import sys

import numpy as np
import pytest

from gpu4gst import GSTGraph, workloads

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import full_size  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(300)
def test_autotune_keeps_results_and_reports(gpu):
    w = workloads.config(2)
    fx = full_size.load(w.name)
    assert fx is not None and fx["digests"] == full_size.digests(w)
    G = GSTGraph.from_generated(w.graph)
    G.set_groups(w.groups)
    single = G.autotune(w.queries[:12], batch_size=1, max_seconds=0)
    many = G.autotune(w.queries, batch_size=0, max_seconds=0)
    assert single["size_class"] == 0 and single["chunk"] == 256
    assert many["size_class"] == 2 and many["chunk"] == 4096
    print(f"tuning the one-query class on 12 queries: {single}", flush=True)
    t1 = G.autotune(w.queries[:12], batch_size=1, max_seconds=20)
    print(f"chosen: {t1}", flush=True)
    assert t1["trials"] >= 3 and t1["seconds_after"] <= t1["seconds_before"]
    assert G.autotune(w.queries[:1], batch_size=1, max_seconds=0)["chunk"] == t1["chunk"]
    for i in range(12):
        assert G.solve(w.queries[i]).cost == fx["costs"][i], i
    print("tuning the 32+ class on 256 queries", flush=True)
    t2 = G.autotune(w.queries[:256], batch_size=0, max_seconds=30)
    print(f"chosen: {t2}", flush=True)
    assert t2["size_class"] == 2 and t2["trials"] >= 3
    res = G.solve_batch(w.queries)
    assert res.costs.tolist() == fx["costs"]
    G.close()
