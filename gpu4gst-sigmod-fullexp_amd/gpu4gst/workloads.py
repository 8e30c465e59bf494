"""The five seeded synthetic workloads the solver is measured and tested on.

 1  R-MAT scale 15 (a,b,c = .57,.19,.19), 262,144 distinct undirected edges,
    largest component; weights U[1,100]; 2,000 groups, each vertex joins
    1 + Poisson(0.5) uniform groups; 100 queries of k = 3; no hop bound.
 2  Chung-Lu V = 168,000, 6.8M distinct undirected edges, degree exponent 2.1,
    largest expected degree 35,000, largest component; weights U[1,1000];
    10,000 groups with Zipf(1.2) sizes clipped to [1,5000]; 1,024 queries with
    k cycling 3, 5, 7 (342 / 341 / 341); no hop bound.
 3  Chung-Lu V = 1.1M, 3.0M edges, exponent 2.3 (largest expected degree
    5,000); weights U[1,100]; 20,000 groups Zipf(1.3) in [1,2000]; 512
    queries of k = 5 with hop bound D uniform in {3,4,5}.
 4  Chung-Lu V = 317,000, 1.05M edges, exponent 2.4 (largest expected degree
    3,000); weights U[1,1000]; 5,000 groups Zipf(1.2) in [1,5000]; 256 queries
    of k = 5, differentiable.
 5  R-MAT scale 22 (V = 4,194,304), 35M distinct edges; weights U[1,1000];
    50,000 groups Zipf(1.2) in [1,20000]; 256 queries of k = 7.

Workloads 1 and 2 keep only the largest component, as their definitions say.
Workloads 3, 4 and 5 keep every vertex (their stated V, so every per-query
table has its full size: 4,194,304 x 128 states per query on workload 5),
and draw group members uniformly from the vertices of the largest component,
so that every query is feasible (on R-MAT scale 22 half of the vertices lie
outside it, most of them isolated).

Streams of seeds: seed(cfg, stream) = 0x67535400 + 16 cfg + stream, stream 0
graph, 1 weights, 2 groups, 3 queries, 4 hop bounds / sampling.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import gen


def seed(cfg: int, stream: int) -> int:
    return 0x67535400 + 16 * cfg + stream


@dataclass
class Workload:
    name: str
    graph: gen.Graph
    groups: gen.Groups
    queries: list
    D: np.ndarray | None  # per-query hop bound, None = plain


def _queries(groups, ks, cfg):
    return gen.queries(groups, np.asarray(ks, np.int32), seed=seed(cfg, 3))


def config(cfg: int, n_queries: int | None = None) -> Workload:
    if cfg == 1:
        g = gen.rmat(15, 262144, 0.57, 0.19, 0.19, seed=seed(1, 0), wlo=1, whi=100,
                     wseed=seed(1, 1))
        grp = gen.groups_poisson(g.n, 2000, 0.5, seed=seed(1, 2))
        nq = n_queries or 100
        return Workload("rmat15_k3", g, grp, _queries(grp, [3] * nq, 1), None)
    if cfg == 2:
        g = gen.chung_lu(168000, 6800000, 2.1, 35000.0, seed=seed(2, 0), wlo=1, whi=1000,
                         wseed=seed(2, 1))
        grp = gen.groups_zipf(g.n, 10000, 1.2, 1, 5000, seed=seed(2, 2))
        nq = n_queries or 1024
        ks = [(3, 5, 7)[i % 3] for i in range(nq)]
        return Workload("chunglu168k_k357", g, grp, _queries(grp, ks, 2), None)
    if cfg == 3:
        g = gen.chung_lu(1100000, 3000000, 2.3, 5000.0, seed=seed(3, 0), wlo=1, whi=100,
                         wseed=seed(3, 1), lcc=False)
        grp = gen.groups_zipf_pool(g.n, g.lcc, 20000, 1.3, 1, 2000, seed=seed(3, 2))
        nq = n_queries or 512
        D = gen.uniform_ints(nq, 3, 5, seed=seed(3, 4))
        return Workload("chunglu1100k_k5_hop", g, grp, _queries(grp, [5] * nq, 3), D)
    if cfg == 4:
        g = gen.chung_lu(317000, 1050000, 2.4, 3000.0, seed=seed(4, 0), wlo=1, whi=1000,
                         wseed=seed(4, 1), lcc=False)
        grp = gen.groups_zipf_pool(g.n, g.lcc, 5000, 1.2, 1, 5000, seed=seed(4, 2))
        nq = n_queries or 256
        return Workload("chunglu317k_k5_grad", g, grp, _queries(grp, [5] * nq, 4), None)
    if cfg == 5:
        g = gen.rmat(22, 35000000, 0.57, 0.19, 0.19, seed=seed(5, 0), wlo=1, whi=1000,
                     wseed=seed(5, 1), lcc=False)
        grp = gen.groups_zipf_pool(g.n, g.lcc, 50000, 1.2, 1, 20000, seed=seed(5, 2))
        nq = n_queries or 256
        return Workload("rmat22_k7", g, grp, _queries(grp, [7] * nq, 5), None)
    raise ValueError(f"unknown workload {cfg}")


def small(n: int, m: int, G: int, *, seed_: int, wlo=1, whi=20, member_p=0.2):
    """A small random connected graph (R-MAT-free, uniform) for parity tests."""
    rng = np.random.default_rng(seed_)
    perm = rng.permutation(n)
    src = np.concatenate([perm[:-1], rng.integers(0, n, m)])
    dst = np.concatenate([perm[1:], rng.integers(0, n, m)])
    w = rng.integers(wlo, whi + 1, src.shape[0])
    g = gen.from_edges(n, src, dst, w)
    lists = []
    for _ in range(G):
        mem = np.nonzero(rng.random(n) < member_p)[0].tolist()
        if not mem:
            mem = [int(rng.integers(0, n))]
        lists.append(mem)
    return g, gen.groups_from_lists(lists)


def degree_sequence_graph(degrees, *, seed_: int, wlo: int = 1, whi: int = 100) -> gen.Graph:
    """Configuration-model graph on a given degree sequence (stubs paired
    uniformly at random; loops and repeated pairs dropped), vertex ids kept,
    weights U[wlo, whi]. Used with the reference's Twitch_beg_pos.bin, whose
    edge and weight files are not shipped, to get a graph of Twitch's exact
    vertex count and (up to the dropped pairs) degree sequence."""
    deg = np.asarray(degrees, np.int64)
    rng = np.random.default_rng(seed_)
    stubs = np.repeat(np.arange(deg.shape[0], dtype=np.int64), deg)
    rng.shuffle(stubs)
    stubs = stubs[: stubs.shape[0] // 2 * 2].reshape(-1, 2)
    keep = stubs[:, 0] != stubs[:, 1]
    src, dst = stubs[keep, 0], stubs[keep, 1]
    w = rng.integers(wlo, whi + 1, src.shape[0])
    return gen.from_edges(int(deg.shape[0]), src, dst, w, lcc=False)
