// Seeded synthetic inputs for group Steiner tree workloads: power-law graphs
// (R-MAT, Chung-Lu), vertex groups (Poisson memberships, Zipf sizes) and
// queries of k distinct groups.
//
// The reference ships no generator on its live path (a random connected
// graph generator exists unused at
// code/D-TrimCDP/include/graph_v_of_v_idealID/random_graph/
// graph_v_of_v_idealID_generate_random_connected_graph.h:7-107), and its only
// dataset has no edge files, so every workload here is produced by this file.
// All randomness comes from splitmix64 with explicit integer mappings, so the
// same seed gives the same graph on every machine that runs this binary.
#include "../../include/gst_gen.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform integer in [0, n)
  uint64_t below(uint64_t n) {
    return (uint64_t)(((unsigned __int128)next() * n) >> 64);
  }
  // uniform double in [0, 1)
  double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Open-addressing set of undirected edge keys (min << 32 | max).
struct EdgeSet {
  std::vector<uint64_t> slots;
  uint64_t mask;
  explicit EdgeSet(int64_t expected) {
    uint64_t cap = 1024;
    while (cap < (uint64_t)expected * 2) cap <<= 1;
    slots.assign(cap, ~0ull);
    mask = cap - 1;
  }
  bool insert(uint64_t key) {
    uint64_t h = key * 0x9E3779B97F4A7C15ull;
    uint64_t i = (h ^ (h >> 29)) & mask;
    while (true) {
      if (slots[i] == ~0ull) { slots[i] = key; return true; }
      if (slots[i] == key) return false;
      i = (i + 1) & mask;
    }
  }
};

}  // namespace

struct gg_graph {
  int64_t n = 0;
  std::vector<int32_t> src, dst, w;  // undirected, src < dst, insertion order
};

struct gg_groups {
  int64_t n = 0;
  std::vector<int64_t> off;
  std::vector<int32_t> mem;
};

static void add_unique_edges(gg_graph* g, EdgeSet& set, int64_t a, int64_t b) {
  if (a == b) return;
  int64_t lo = std::min(a, b), hi = std::max(a, b);
  if (set.insert(((uint64_t)lo << 32) | (uint64_t)hi)) {
    g->src.push_back((int32_t)lo);
    g->dst.push_back((int32_t)hi);
  }
}

extern "C" {

gg_graph* gg_rmat(int32_t scale, int64_t m_target, double a, double b, double c,
                  uint64_t seed) {
  if (scale < 1 || scale > 30 || m_target < 0) return nullptr;
  gg_graph* g = new gg_graph;
  g->n = (int64_t)1 << scale;
  int64_t max_edges = g->n * (g->n - 1) / 2;
  if (m_target > max_edges) m_target = max_edges;
  EdgeSet set(m_target);
  Rng rng(seed);
  double ab = a + b, abc = a + b + c;
  while ((int64_t)g->src.size() < m_target) {
    int64_t u = 0, v = 0;
    for (int l = 0; l < scale; ++l) {
      double r = rng.unit();
      int bu = 0, bv = 0;
      if (r < a) {
      } else if (r < ab) {
        bv = 1;
      } else if (r < abc) {
        bu = 1;
      } else {
        bu = 1;
        bv = 1;
      }
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    add_unique_edges(g, set, u, v);
  }
  return g;
}

gg_graph* gg_chung_lu(int64_t n, int64_t m_target, double beta, double dmax,
                      uint64_t seed) {
  if (n < 2 || m_target < 0 || beta <= 1.0 || dmax <= 0) return nullptr;
  // expected degree of vertex i is A * (i + i0)^(-alpha), alpha = 1/(beta-1);
  // i0 is found by bisection so that the largest expected degree is dmax
  // while the expected degrees sum to 2 * m_target.
  const double alpha = 1.0 / (beta - 1.0);
  const double total = 2.0 * (double)m_target;
  auto ratio = [&](double i0) {  // max / sum for a given offset
    double s = 0;
    for (int64_t i = 0; i < n; ++i) s += std::pow((double)i + i0, -alpha);
    return std::pow(i0, -alpha) / s;
  };
  double want = dmax / total, lo = 1e-6, hi = (double)n * 1e3;
  for (int it = 0; it < 64; ++it) {
    double mid = std::sqrt(lo * hi);
    if (ratio(mid) > want) lo = mid; else hi = mid;
  }
  const double i0 = std::sqrt(lo * hi);
  std::vector<double> cdf(n);
  double acc = 0;
  for (int64_t i = 0; i < n; ++i) {
    acc += std::pow((double)i + i0, -alpha);
    cdf[i] = acc;
  }
  gg_graph* g = new gg_graph;
  g->n = n;
  int64_t max_edges = n * (n - 1) / 2;
  if (m_target > max_edges) m_target = max_edges;
  EdgeSet set(m_target);
  Rng rng(seed);
  auto draw = [&]() {
    double r = rng.unit() * acc;
    int64_t i = std::upper_bound(cdf.begin(), cdf.end(), r) - cdf.begin();
    return i < n ? i : n - 1;
  };
  while ((int64_t)g->src.size() < m_target) {
    int64_t u = draw(), v = draw();
    add_unique_edges(g, set, u, v);
  }
  return g;
}

gg_graph* gg_from_edges(int64_t n, int64_t m, const int32_t* src,
                        const int32_t* dst, const int32_t* w) {
  if (n < 1 || m < 0) return nullptr;
  gg_graph* g = new gg_graph;
  g->n = n;
  EdgeSet set(m + 1);
  for (int64_t i = 0; i < m; ++i) {
    if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n) {
      delete g;
      return nullptr;
    }
    size_t before = g->src.size();
    add_unique_edges(g, set, src[i], dst[i]);
    if (w && g->src.size() > before) g->w.push_back(w[i]);
  }
  return g;
}

void gg_keep_lcc(gg_graph* g) {
  std::vector<int64_t> parent(g->n);
  std::iota(parent.begin(), parent.end(), 0);
  auto find = [&](int64_t x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  for (size_t i = 0; i < g->src.size(); ++i) {
    int64_t a = find(g->src[i]), b = find(g->dst[i]);
    if (a != b) parent[std::max(a, b)] = std::min(a, b);
  }
  std::vector<int64_t> size(g->n, 0);
  for (int64_t v = 0; v < g->n; ++v) size[find(v)]++;
  int64_t best = 0;
  for (int64_t v = 1; v < g->n; ++v)
    if (size[v] > size[best]) best = v;
  std::vector<int32_t> relabel(g->n, -1);
  int64_t nn = 0;
  for (int64_t v = 0; v < g->n; ++v)
    if (find(v) == best) relabel[v] = (int32_t)nn++;
  size_t k = 0;
  for (size_t i = 0; i < g->src.size(); ++i) {
    int32_t a = relabel[g->src[i]], b = relabel[g->dst[i]];
    if (a < 0 || b < 0) continue;
    g->src[k] = a;
    g->dst[k] = b;
    if (!g->w.empty()) g->w[k] = g->w[i];
    ++k;
  }
  g->src.resize(k);
  g->dst.resize(k);
  if (!g->w.empty()) g->w.resize(k);
  g->n = nn;
}

int64_t gg_lcc_vertices(const gg_graph* g, int32_t* out) {
  std::vector<int64_t> parent(g->n);
  std::iota(parent.begin(), parent.end(), 0);
  auto find = [&](int64_t x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  for (size_t i = 0; i < g->src.size(); ++i) {
    int64_t a = find(g->src[i]), b = find(g->dst[i]);
    if (a != b) parent[std::max(a, b)] = std::min(a, b);
  }
  std::vector<int64_t> size(g->n, 0);
  for (int64_t v = 0; v < g->n; ++v) size[find(v)]++;
  int64_t best = 0;
  for (int64_t v = 1; v < g->n; ++v)
    if (size[v] > size[best]) best = v;
  int64_t k = 0;
  for (int64_t v = 0; v < g->n; ++v)
    if (find(v) == best) {
      if (out) out[k] = (int32_t)v;
      ++k;
    }
  return k;
}

void gg_set_weights_uniform(gg_graph* g, int32_t lo, int32_t hi, uint64_t seed) {
  Rng rng(seed);
  g->w.resize(g->src.size());
  uint64_t span = (uint64_t)((int64_t)hi - lo + 1);
  for (auto& x : g->w) x = lo + (int32_t)rng.below(span);
}

int64_t gg_num_vertices(const gg_graph* g) { return g->n; }
int64_t gg_num_edges(const gg_graph* g) { return (int64_t)g->src.size(); }

void gg_get_edges(const gg_graph* g, int32_t* src, int32_t* dst, int32_t* w) {
  size_t m = g->src.size();
  std::memcpy(src, g->src.data(), m * 4);
  std::memcpy(dst, g->dst.data(), m * 4);
  if (w) {
    if (g->w.size() == m) std::memcpy(w, g->w.data(), m * 4);
    else std::fill(w, w + m, 1);
  }
}

void gg_get_csr(const gg_graph* g, int64_t* beg, int32_t* col, int32_t* w,
                int32_t* eid) {
  const int64_t n = g->n;
  const size_t m = g->src.size();
  std::vector<int64_t> deg(n + 1, 0);
  for (size_t i = 0; i < m; ++i) {
    deg[g->src[i]]++;
    deg[g->dst[i]]++;
  }
  beg[0] = 0;
  for (int64_t v = 0; v < n; ++v) beg[v + 1] = beg[v] + deg[v];
  std::vector<int64_t> pos(beg, beg + n);
  // fill, then sort every row by neighbour id (rows are stored sorted, as the
  // reference loader does at graph_hash_of_mixed_weighted_read_for_GSTP.h:130-133)
  std::vector<int64_t> tmp_eid(2 * m);
  std::vector<int32_t> tmp_col(2 * m);
  for (size_t i = 0; i < m; ++i) {
    int64_t a = pos[g->src[i]]++, b = pos[g->dst[i]]++;
    tmp_col[a] = g->dst[i];
    tmp_eid[a] = (int64_t)i;
    tmp_col[b] = g->src[i];
    tmp_eid[b] = (int64_t)i;
  }
  std::vector<int64_t> idx;
  for (int64_t v = 0; v < n; ++v) {
    int64_t s = beg[v], e = beg[v + 1];
    idx.resize(e - s);
    std::iota(idx.begin(), idx.end(), s);
    std::sort(idx.begin(), idx.end(),
              [&](int64_t x, int64_t y) { return tmp_col[x] < tmp_col[y]; });
    for (int64_t j = s; j < e; ++j) {
      int64_t from = idx[j - s];
      col[j] = tmp_col[from];
      int64_t id = tmp_eid[from];
      if (w) w[j] = g->w.size() == m ? g->w[id] : 1;
      if (eid) eid[j] = (int32_t)id;
    }
  }
}

void gg_free(gg_graph* g) { delete g; }

gg_groups* gg_groups_poisson(int64_t n, int32_t G, double lambda, uint64_t seed) {
  if (n < 1 || G < 1 || lambda < 0) return nullptr;
  Rng rng(seed);
  std::vector<std::vector<int32_t>> members(G);
  const double L = std::exp(-lambda);
  std::vector<int32_t> chosen;
  for (int64_t v = 0; v < n; ++v) {
    // Knuth's multiplication method for Poisson(lambda)
    int extra = 0;
    double p = rng.unit();
    while (p > L) {
      ++extra;
      p *= rng.unit();
    }
    int want = std::min<int64_t>(1 + extra, G);
    chosen.clear();
    while ((int)chosen.size() < want) {
      int32_t gid = (int32_t)rng.below(G);
      if (std::find(chosen.begin(), chosen.end(), gid) == chosen.end())
        chosen.push_back(gid);
    }
    for (int32_t gid : chosen) members[gid].push_back((int32_t)v);
  }
  gg_groups* out = new gg_groups;
  out->n = n;
  out->off.assign(G + 1, 0);
  for (int32_t i = 0; i < G; ++i) out->off[i + 1] = out->off[i] + members[i].size();
  out->mem.reserve(out->off[G]);
  for (auto& m : members) out->mem.insert(out->mem.end(), m.begin(), m.end());
  return out;
}

gg_groups* gg_groups_zipf(int64_t n, int32_t G, double s, int32_t smin,
                          int32_t smax, uint64_t seed) {
  return gg_groups_zipf_pool(n, nullptr, n, G, s, smin, smax, seed);
}

gg_groups* gg_groups_zipf_pool(int64_t n, const int32_t* pool, int64_t npool, int32_t G,
                               double s, int32_t smin, int32_t smax, uint64_t seed) {
  if (n < 1 || npool < 1 || G < 1 || smin < 1 || smax < smin) return nullptr;
  if (pool)
    for (int64_t i = 0; i < npool; ++i)
      if (pool[i] < 0 || pool[i] >= n) return nullptr;
  const int64_t np = pool ? npool : n;
  if (smax > np) smax = (int32_t)np;
  if (smin > smax) smin = smax;
  Rng rng(seed);
  // truncated Zipf over sizes 1..smax, then clipped to [smin, smax]
  std::vector<double> cdf(smax);
  double acc = 0;
  for (int32_t x = 1; x <= smax; ++x) {
    acc += std::pow((double)x, -s);
    cdf[x - 1] = acc;
  }
  gg_groups* out = new gg_groups;
  out->n = n;
  out->off.assign(G + 1, 0);
  std::vector<int32_t> mem;
  std::vector<int32_t> picked;
  for (int32_t gi = 0; gi < G; ++gi) {
    double r = rng.unit() * acc;
    int32_t size = (int32_t)(std::upper_bound(cdf.begin(), cdf.end(), r) - cdf.begin()) + 1;
    size = std::max(smin, std::min(smax, size));
    // Floyd's algorithm: `size` distinct uniform positions of the pool
    picked.clear();
    EdgeSet seen(size + 1);
    for (int64_t j = np - size; j < np; ++j) {
      int64_t t = (int64_t)rng.below((uint64_t)j + 1);
      if (seen.insert((uint64_t)t)) picked.push_back((int32_t)t);
      else {
        seen.insert((uint64_t)j);
        picked.push_back((int32_t)j);
      }
    }
    if (pool)
      for (auto& x : picked) x = pool[x];
    std::sort(picked.begin(), picked.end());
    out->off[gi + 1] = out->off[gi] + size;
    mem.insert(mem.end(), picked.begin(), picked.end());
  }
  out->mem = std::move(mem);
  return out;
}

int64_t gg_groups_count(const gg_groups* g) { return (int64_t)g->off.size() - 1; }
int64_t gg_groups_total(const gg_groups* g) { return (int64_t)g->mem.size(); }
void gg_groups_get(const gg_groups* g, int64_t* off, int32_t* mem) {
  std::memcpy(off, g->off.data(), g->off.size() * 8);
  std::memcpy(mem, g->mem.data(), g->mem.size() * 4);
}
void gg_groups_free(gg_groups* g) { delete g; }

int32_t gg_queries(int32_t G, const int64_t* goff, int32_t nq, const int32_t* ks,
                   uint64_t seed, int32_t* out_groups) {
  Rng rng(seed);
  int64_t w = 0;
  int32_t nonempty = 0;
  for (int32_t i = 0; i < G; ++i) nonempty += goff[i + 1] > goff[i];
  for (int32_t q = 0; q < nq; ++q) {
    int32_t k = ks[q];
    if (k < 1 || k > nonempty) return -1;
    int64_t start = w;
    while (w - start < k) {
      int32_t gid = (int32_t)rng.below(G);
      if (goff[gid + 1] == goff[gid]) continue;  // queries use non-empty groups
      bool dup = false;
      for (int64_t j = start; j < w; ++j) dup |= out_groups[j] == gid;
      if (!dup) out_groups[w++] = gid;
    }
  }
  return 0;
}

void gg_uniform_ints(int64_t count, int32_t lo, int32_t hi, uint64_t seed,
                     int32_t* out) {
  Rng rng(seed);
  uint64_t span = (uint64_t)((int64_t)hi - lo + 1);
  for (int64_t i = 0; i < count; ++i) out[i] = lo + (int32_t)rng.below(span);
}

}  // extern "C"
