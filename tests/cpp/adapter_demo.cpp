// Calls the reference-shaped adapters of include/gst_reference_adapters.hpp
// with stand-in types laid out like the reference's (CSR_graph fields,
// records counters, node, non_overlapped_group_sets). Prints one JSON object
// per query: {"D": .., "cost": .., "status": .., "edges": [[u, v], ...]}.
// Input on stdin: V M, M lines "u v w", G, G lines "n v1 .. vn", Q, Q lines "D k g1 .. gk".
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "gst_reference_adapters.hpp"

struct csr_stand_in {  // the reference's CSR_graph exposes these pointers
  int *all_pointer = nullptr, *all_edge = nullptr, *all_edge_weight = nullptr;
  int V = 0;
};
struct records_stand_in {
  int process_queue_num = 0, counts = 0, mid_process_queue_num = 0, mid_counts = 0;
  double mid_time = 0;
};
struct node_stand_in {
  int update = 0, type = 0, cost = 0, c = 0, u = 0, p1 = 0, p2 = 0;
};
struct nogs_stand_in {
  std::vector<int> pointer, ids;
  int length = 0;
};

int selftest() {
  // 4-cycle 0-1 (5), 1-2 (1), 2-3 (1), 3-0 (1); groups {0} and {3}, {1}
  gst_ref::adjacency g(4), groups(3);
  auto add = [&](int a, int b, int w) {
    g[a].push_back({b, w});
    g[b].push_back({a, w});
  };
  add(0, 1, 5);
  add(1, 2, 1);
  add(2, 3, 1);
  add(3, 0, 1);
  groups[0] = {{0, 1}};
  groups[1] = {{3, 1}};
  groups[2] = {{1, 1}};
  std::vector<int> q{0, 1, 2};
  gst_ref::tree path;
  path.edges = {{0, 1}, {1, 2}, {2, 3}};
  std::string why;
  std::printf("selftest path_feasible %d\n", (int)gst_ref::feasible(path, groups, q, -1, &why));
  std::printf("selftest path_cost %lld\n", gst_ref::tree_cost(path, g));
  std::printf("selftest path_mst_cost %d\n", gst_ref::mst_postprocess(path, g).cost);
  gst_ref::tree split;
  split.edges = {{0, 1}, {2, 3}};
  std::printf("selftest split_feasible %d\n", (int)gst_ref::feasible(split, groups, q, -1, &why));
  std::printf("selftest split_why %s\n", why.find("disconnected") != std::string::npos ? "disconnected" : "other");
  gst_ref::tree part;
  part.edges = {{1, 2}, {2, 3}};
  std::printf("selftest part_feasible %d\n", (int)gst_ref::feasible(part, groups, q, -1, &why));
  std::printf("selftest part_why %s\n", why.find("not covered") != std::string::npos ? "uncovered" : "other");
  gst_ref::tree single;
  std::vector<int> q1{2};
  std::printf("selftest single_feasible %d\n", (int)gst_ref::feasible(single, groups, q1, 1, &why));
  return 0;
}

int main(int argc, char** argv) {
  if (argc > 1 && std::string(argv[1]) == "selftest") return selftest();
  int V, M;
  std::cin >> V >> M;
  gst_ref::adjacency input_graph(V);
  for (int i = 0; i < M; ++i) {
    int a, b, w;
    std::cin >> a >> b >> w;
    input_graph[a].push_back({b, w});
    input_graph[b].push_back({a, w});
  }
  int G;
  std::cin >> G;
  gst_ref::adjacency group_graph(G);
  for (int g = 0; g < G; ++g) {
    int n;
    std::cin >> n;
    for (int j = 0; j < n; ++j) {
      int v;
      std::cin >> v;
      group_graph[g].push_back({v, 1});
    }
  }
  int Q;
  std::cin >> Q;
  csr_stand_in csr;
  csr.V = V;
  for (int q = 0; q < Q; ++q) {
    int D, k;
    std::cin >> D >> k;
    std::vector<int> query(k);
    for (auto& x : query) std::cin >> x;
    records_stand_in ret;
    gst_ref::tree t;
    int cost = -1;
    if (D < 0) {
      double rt = 0;
      int ram = 0;
      t = gst_ref::dp_gpu((node_stand_in**)nullptr, (node_stand_in*)nullptr, csr, query,
                          group_graph, input_graph, cost, nogs_stand_in{}, rt, ram, ret, 0);
    } else {
      std::vector<unsigned int> uq(query.begin(), query.end());
      double rt = 0;
      long long ram = 0;
      t = gst_ref::dp_gpu_hop(csr, uq, group_graph, input_graph, D, &rt, cost, ram, ret, 0);
    }
    // a tree without edges is a vertex in every query group
    int root = -1;
    for (int v = 0; v < V && root < 0 && t.edges.empty(); ++v) {
      bool all = true;
      for (int gq : query) {
        bool in = false;
        for (const auto& m : group_graph[gq]) in |= m.first == v;
        all &= in;
      }
      if (all) root = v;
    }
    const int ok = t.status == GST_Q_OK ? (int)gst_ref::feasible(t, group_graph, query, root) : 0;
    std::printf("{\"D\": %d, \"cost\": %d, \"status\": %d, \"feasible\": %d, "
                "\"tree_cost\": %lld, \"mst_cost\": %d, \"edges\": [",
                D, cost, t.status, ok, gst_ref::tree_cost(t, input_graph),
                gst_ref::mst_postprocess(t, input_graph).cost);
    for (size_t i = 0; i < t.edges.size(); ++i)
      std::printf("%s[%d, %d]", i ? ", " : "", t.edges[i].first, t.edges[i].second);
    std::printf("]}\n");
  }
  return 0;
}
