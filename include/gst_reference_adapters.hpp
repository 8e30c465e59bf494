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
// (INTEGRATION.md). Unlike the GPU reference they also return the tree.
#ifndef GST_REFERENCE_ADAPTERS_HPP
#define GST_REFERENCE_ADAPTERS_HPP

#include <chrono>
#include <cstdint>
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

}  // namespace gst_ref

#endif
