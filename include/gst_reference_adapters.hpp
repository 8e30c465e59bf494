// gst_reference_adapters.hpp - C++ adapters with the reference solvers' call shapes.
//
// Header-only, over the C ABI of gst.h. The reference's harnesses call
//   DP_GPU(host_tree, host_tree_one_d, graph, query, group_graph, input_graph, real_cost,
//          non_overlapped_group_sets, rt, RAM, ret, res_weight)
//          (code/TrimCDP/include/DPQ.cuh:67, src/DPQ.cu:292)
//   DP_gpu(graph, query, group_graph, input_graph, D, &rt, real_cost, RAM, ret, res_weight)
//          (code/D-TrimCDP/include/DPQ.cuh:66, src/DPQ.cu:147)
// with the adjacency-list containers of the reference
// (graph_v_of_v_idealID = std::vector<std::vector<std::pair<int, int>>>,
// code/TrimCDP/include/graph_v_of_v_idealID/graph_v_of_v_idealID.h:50).
// gst_ref::dp_gpu / gst_ref::dp_gpu_hop take the same arguments (the
// reference's own CSR_graph / records / node types bind to the template
// parameters) and fill the same outputs: real_cost, rt (seconds of the solve),
// RAM (finite-state count stand-in: DP states expanded) and ret's counters.
// A harness keeps its types and defines DP_GPU / DP_gpu as one-line forwards
// to these. Unlike the GPU reference they also return the tree.
#ifndef GST_REFERENCE_ADAPTERS_HPP
#define GST_REFERENCE_ADAPTERS_HPP

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <functional>
#include <map>
#include <queue>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gst.h"

namespace gst_ref {

typedef std::vector<std::vector<std::pair<int, int>>> adjacency;

// One device graph per (input graph, group graph) pair, built on first use and
// rebuilt when the caller passes different containers.
struct context {
  gst_graph* g = nullptr;
  const void* input_key = nullptr;
  const void* group_key = nullptr;
  size_t n_vertices = 0, n_groups = 0;
  ~context() { gst_graph_destroy(g); }
};

inline context& ctx() {
  static context c;
  return c;
}

inline void check(int rc) {
  if (rc != GST_OK) throw std::runtime_error(std::string("gst: ") + gst_last_error());
}

// The reference's input_graph holds both directions of every edge
// (read_input_graph, graph_hash_of_mixed_weighted_read_for_GSTP.h:75-148),
// i.e. it already is a symmetric adjacency; group_graph[g] lists the members
// of group g as (vertex, 1) pairs (read_Group, same file:190-239).
inline gst_graph* graph_for(const adjacency& input_graph, const adjacency& group_graph) {
  context& c = ctx();
  if (c.g && c.input_key == &input_graph && c.group_key == &group_graph &&
      c.n_vertices == input_graph.size() && c.n_groups == group_graph.size())
    return c.g;
  gst_graph_destroy(c.g);
  c.g = nullptr;
  const int64_t V = (int64_t)input_graph.size();
  std::vector<int64_t> beg(V + 1, 0), col, w;
  for (int64_t v = 0; v < V; ++v) {
    beg[v + 1] = beg[v] + (int64_t)input_graph[v].size();
    for (const auto& e : input_graph[v]) {
      col.push_back(e.first);
      w.push_back(e.second);
    }
  }
  check(gst_graph_create(beg.data(), col.data(), w.data(), V, beg[V], 64, &c.g));
  std::vector<int64_t> off(group_graph.size() + 1, 0);
  std::vector<int32_t> mem;
  for (size_t i = 0; i < group_graph.size(); ++i) {
    off[i + 1] = off[i] + (int64_t)group_graph[i].size();
    for (const auto& m : group_graph[i]) mem.push_back(m.first);
  }
  check(gst_groups_set(c.g, off.data(), mem.data(), (int32_t)group_graph.size()));
  c.input_key = &input_graph;
  c.group_key = &group_graph;
  c.n_vertices = input_graph.size();
  c.n_groups = group_graph.size();
  return c.g;
}

struct tree {
  int32_t cost = 0;
  int32_t status = GST_Q_OK;
  std::vector<std::pair<int, int>> edges;
};

template <class Q>
tree solve(const adjacency& input_graph, const adjacency& group_graph, const std::vector<Q>& query,
           int D, double& seconds, gst_stats& st) {
  gst_graph* g = graph_for(input_graph, group_graph);
  std::vector<int32_t> q(query.begin(), query.end());
  std::vector<int32_t> uv(2 * (input_graph.size() + 1));
  int32_t ne = 0;
  tree t;
  const auto t0 = std::chrono::high_resolution_clock::now();
  check(gst_solve(g, q.data(), (int32_t)q.size(), D, &t.cost, uv.data(),
                  (int32_t)(uv.size() / 2), &ne, &t.status, &st));
  seconds = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
  for (int32_t i = 0; i < ne; ++i) t.edges.push_back({uv[2 * i], uv[2 * i + 1]});
  return t;
}

// Plain TrimCDP, the shape of DP_GPU. Infeasible queries report 0, as
// DP_GPU does (DPQ.cu:606-610).
template <class Node, class CSR, class NonOverlapped, class Records>
tree dp_gpu(Node** /*host_tree*/, Node* /*host_tree_one_d*/, CSR& /*graph*/,
            std::vector<int>& query, adjacency& group_graph, adjacency& input_graph,
            int& real_cost, NonOverlapped /*s*/, double& rt, int& RAM, Records& ret,
            int /*res_weight*/) {
  gst_stats st{};
  tree t = solve(input_graph, group_graph, query, -1, rt, st);
  real_cost = t.status == GST_Q_OK ? t.cost : 0;
  RAM = (int)st.states_expanded;
  ret.process_queue_num = (int)st.items_processed;
  ret.counts = (int)st.improvements;
  return t;
}

// Hop-bounded D-TrimCDP, the shape of DP_gpu (which returns an empty tree;
// this one returns the optimal tree of hop diameter <= D).
template <class CSR, class Records>
tree dp_gpu_hop(CSR& /*graph*/, std::vector<unsigned int>& query, adjacency& group_graph,
                adjacency& input_graph, int D, double* rt, int& real_cost, long long& RAM,
                Records& ret, int /*res_weight*/) {
  gst_stats st{};
  double sec = 0;
  tree t = solve(input_graph, group_graph, query, D, sec, st);
  if (rt) *rt = sec;
  real_cost = t.status == GST_Q_OK ? t.cost : 100000;
  RAM = (long long)st.states_expanded;
  ret.process_queue_num = (int)st.items_processed;
  ret.counts = (int)st.improvements;
  return t;
}

// Weight of the edge (a, b) in an adjacency (the smallest of parallel
// copies), or -1 when there is none.
inline int edge_weight(const adjacency& g, int a, int b) {
  int w = -1;
  if (a < 0 || (size_t)a >= g.size()) return -1;
  for (const auto& e : g[a])
    if (e.first == b && (w < 0 || e.second < w)) w = e.second;
  return w;
}

// Sum of the tree's edge weights (graph_hash_of_mixed_weighted_sum_of_ec,
// the cost the reference's CPU harness reports, code/PrunedDP++/include/
// exp_CPU_nonHop.h:117); -1 if an edge is not in the graph.
inline long long tree_cost(const tree& t, const adjacency& input_graph) {
  long long c = 0;
  for (const auto& e : t.edges) {
    const int w = edge_weight(input_graph, e.first, e.second);
    if (w < 0) return -1;
    c += w;
  }
  return c;
}

// The checks of this_is_a_feasible_solution_gpu (code/PrunedDP++/include/
// exp_CPU_nonHop.h:17-49): the tree's vertices form one connected component
// and every query group has a member among them. A tree without edges is the
// single vertex `root` (pass the vertex that covers every group; -1 = none).
// On failure *why holds the reference's message.
template <class Q>
bool feasible(const tree& t, const adjacency& group_graph, const std::vector<Q>& query,
              int root = -1, std::string* why = nullptr) {
  std::map<int, std::vector<int>> adj;
  for (const auto& e : t.edges) {
    adj[e.first].push_back(e.second);
    adj[e.second].push_back(e.first);
  }
  if (adj.empty() && root >= 0) adj[root];
  if (adj.empty()) {
    if (why) *why = "this_is_a_feasible_solution: solu is disconnected!";
    return false;
  }
  std::set<int> seen{adj.begin()->first};
  std::vector<int> st{adj.begin()->first};
  while (!st.empty()) {
    const int v = st.back();
    st.pop_back();
    for (int u : adj[v])
      if (seen.insert(u).second) st.push_back(u);
  }
  if (seen.size() != adj.size()) {
    if (why) *why = "this_is_a_feasible_solution: solu is disconnected!";
    return false;
  }
  for (const auto& gq : query) {
    const size_t gi = (size_t)gq;
    bool covered = false;
    if (gi < group_graph.size())
      for (const auto& m : group_graph[gi])
        if (seen.count(m.first)) {
          covered = true;
          break;
        }
    if (!covered) {
      if (why) *why = "this_is_a_feasible_solution: a group is not covered!";
      return false;
    }
  }
  return true;
}

// MST post-processing (graph_hash_of_mixed_weighted_MST_postprocessing,
// code/PrunedDP++/include/graph_hash_of_mixed_weighted_MST_postprocessing.h:
// 38-66): the minimum spanning tree of the subgraph of input_graph induced by
// the tree's vertices (Prim from the smallest vertex id). It never costs more
// than the tree; for an optimal tree it costs the same.
inline tree mst_postprocess(const tree& t, const adjacency& input_graph) {
  std::set<int> verts;
  for (const auto& e : t.edges) {
    verts.insert(e.first);
    verts.insert(e.second);
  }
  tree out;
  out.status = t.status;
  if (verts.empty()) return out;
  typedef std::pair<long long, std::pair<int, int>> item;  // (w, (to, from))
  std::priority_queue<item, std::vector<item>, std::greater<item>> pq;
  std::set<int> in;
  const int start = *verts.begin();
  pq.push({0, {start, start}});
  long long cost = 0;
  while (!pq.empty() && in.size() < verts.size()) {
    const item x = pq.top();
    pq.pop();
    const int v = x.second.first;
    if (in.count(v)) continue;
    in.insert(v);
    if (v != x.second.second) {
      out.edges.push_back({x.second.second, v});
      cost += x.first;
    }
    for (const auto& e : input_graph[v])
      if (verts.count(e.first) && !in.count(e.first)) pq.push({e.second, {e.first, v}});
  }
  out.cost = (int32_t)cost;
  return out;
}

}  // namespace gst_ref

#endif
