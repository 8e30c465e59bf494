// gst_run - the reference harnesses' command lines over libgst.so.
//
//   gst_run TrimCDP      2 <path> <data_name> <T> <first> <last>
//   gst_run D-TrimCDP    2 <path> <data_name> <T> <D> <first> <last>
//   gst_run TrimCDP-WB   2 <path> <data_name> <T> <first> <last>
//   gst_run D-TrimCDP-WB 2 <path> <data_name> <T> <D> <first> <last>
//   [--single] [--chunk N]
//
// The argument lists are those of code/TrimCDP/src/main.cpp:14-17,
// code/D-TrimCDP/src/main.cpp:18-21, code/TrimCDP-WB/src/GSTnonHop.cu:162-163 and
// code/D-TrimCDP-WB/src/GPUHop.cu:98-99 (argv[1] = 2 selects the GPU
// experiment). When the program is started under one of the four variant
// names (a symlink), the variant word is left out, so existing scripts run
// unchanged. Inputs: <path><data_name>.in (TrimCDP, D-TrimCDP) or the int64
// <path><data_name>_beg_pos.bin/_csr.bin/_weight.bin (the -WB variants), plus
// <path><data_name>.g and <path><data_name><T>.csv; queries first..last
// (inclusive) are solved. Output: <path>result/<variant>_<data_name>_T<T>_
// [D<D>_]<first>-<last>.csv with the reference's header and row layout.
//
// By default the queries go to gst_solve_batch in one call; the time column
// is then the batch's wall time divided by the query count, the memory
// column the per-query device footprint in MB and the counter columns 0.
// --single solves one query per gst_solve call and fills the time column
// with that call's wall time and the counters with its statistics.
// Infeasible queries get the reference's cost: 0 for TrimCDP
// (code/TrimCDP/src/DPQ.cu:606-610), 100000 otherwise.
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gst.h"
#include "gst_formats.hpp"

namespace {

int usage() {
  std::fprintf(stderr,
               "usage: gst_run TrimCDP|TrimCDP-WB 2 <path> <data_name> <T> <first> <last>\n"
               "       gst_run D-TrimCDP|D-TrimCDP-WB 2 <path> <data_name> <T> <D> <first> <last>\n"
               "       options: --single  --chunk N\n");
  return 2;
}

int die(const std::string& msg) {
  std::fprintf(stderr, "gst_run: %s\n", msg.c_str());
  return 1;
}

bool known(const std::string& v) {
  return v == "TrimCDP" || v == "D-TrimCDP" || v == "TrimCDP-WB" || v == "D-TrimCDP-WB";
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> a;
  bool single = false;
  long long chunk = 0;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--single")) {
      single = true;
    } else if (!std::strcmp(argv[i], "--chunk") && i + 1 < argc) {
      chunk = std::atoll(argv[++i]);
    } else {
      a.push_back(argv[i]);
    }
  }
  std::string variant;
  {
    std::string self = argv[0];
    const size_t slash = self.find_last_of('/');
    if (slash != std::string::npos) self = self.substr(slash + 1);
    if (known(self)) {
      variant = self;
    } else if (!a.empty() && known(a[0])) {
      variant = a[0];
      a.erase(a.begin());
    } else {
      return usage();
    }
  }
  const bool hop = variant[0] == 'D';
  const bool wb = variant.size() > 3 && variant.compare(variant.size() - 3, 3, "-WB") == 0;
  if (a.size() != (hop ? 7u : 6u) || a[0] != "2") return usage();
  const std::string path = a[1], data = a[2];
  const int T = std::atoi(a[3].c_str());
  const int D = hop ? std::atoi(a[4].c_str()) : -1;
  const int first = std::atoi(a[hop ? 5 : 4].c_str());
  const int last = std::atoi(a[hop ? 6 : 5].c_str());
  std::string err;

  gst_fmt::csr64 csr;
  if (wb) {
    if (!gst_fmt::read_bin(path + data, csr, &err)) return die(err);
  } else {
    gst_fmt::adjacency adj;
    if (!gst_fmt::read_in(path + data + ".in", adj, &err)) return die(err);
    csr = gst_fmt::csr_from_adjacency(adj);
  }
  gst_fmt::adjacency groups;
  if (!gst_fmt::read_groups(path + data + ".g", groups, &err)) return die(err);
  std::vector<std::vector<int>> queries;
  if (!gst_fmt::read_queries(path + data + std::to_string(T) + ".csv", queries, &err))
    return die(err);
  if (first < 0 || last < first || last >= (int)queries.size())
    return die("query range " + std::to_string(first) + ".." + std::to_string(last) +
               " is outside the " + std::to_string(queries.size()) + " queries of the file");

  gst_graph* g = nullptr;
  if (gst_graph_create(csr.beg.data(), csr.col.data(), csr.w.data(), csr.vertices(),
                       csr.entries(), 64, &g) != GST_OK)
    return die(gst_last_error());
  std::vector<int64_t> goff;
  std::vector<int32_t> gmem;
  gst_fmt::groups_csr(groups, goff, gmem);
  if (gst_groups_set(g, goff.data(), gmem.data(), (int32_t)groups.size()) != GST_OK) {
    const std::string m = gst_last_error();
    gst_graph_destroy(g);
    return die(m);
  }
  const int32_t infeasible = variant == "TrimCDP" ? 0 : 100000;
  gst_set_option(g, GST_OPT_INFEASIBLE_COST, infeasible);
  if (chunk) gst_set_option(g, GST_OPT_CHUNK, chunk);

  const int n = last - first + 1;
  std::vector<int32_t> costs(n), status(n);
  std::vector<double> seconds(n);
  std::vector<long long> memory(n), counts(n), processed(n);
  int rc = GST_OK;
  if (single) {
    for (int i = 0; i < n && rc == GST_OK; ++i) {
      const auto& q = queries[first + i];
      std::vector<int32_t> qg(q.begin(), q.end());
      gst_stats st{};
      int32_t ne = 0;
      const auto t0 = std::chrono::steady_clock::now();
      rc = gst_solve(g, qg.data(), (int32_t)qg.size(), D, &costs[i], nullptr, 0, &ne, &status[i],
                     &st);
      seconds[i] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      memory[i] = st.states_expanded;
      counts[i] = st.improvements;
      processed[i] = st.items_processed;
    }
  } else {
    std::vector<int64_t> qoff(n + 1, 0);
    std::vector<int32_t> qg, qd(n, D);
    for (int i = 0; i < n; ++i) {
      const auto& q = queries[first + i];
      qg.insert(qg.end(), q.begin(), q.end());
      qoff[i + 1] = (int64_t)qg.size();
    }
    std::vector<int64_t> eoff(n + 1);
    gst_stats st{};
    const auto t0 = std::chrono::steady_clock::now();
    rc = gst_solve_batch(g, qoff.data(), qg.data(), qd.data(), n, 0, costs.data(), status.data(),
                         eoff.data(), nullptr, nullptr, 0, &st);
    if (rc == GST_ERR_CAPACITY) rc = GST_OK;  // trees were not asked for
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < n; ++i) {
      seconds[i] = el / n;
      memory[i] = gst_query_footprint(g, (int32_t)queries[first + i].size(), D) >> 20;
    }
  }
  if (rc != GST_OK) {
    const std::string m = gst_last_error();
    gst_graph_destroy(g);
    return die(m);
  }
  gst_graph_destroy(g);

  mkdir((path + "result").c_str(), 0777);
  const std::string out =
      path + "result/" + gst_fmt::result_file_name(variant, data, T, D, first, last);
  FILE* f = std::fopen(out.c_str(), "w");
  if (!f) return die("cannot create " + out);
  std::fprintf(f, "%s\n", gst_fmt::result_header(variant).c_str());
  for (int i = 0; i < n; ++i)
    std::fprintf(f, "%s\n",
                 gst_fmt::result_row(variant, first + i, queries[first + i], seconds[i], costs[i],
                                     memory[i], counts[i], processed[i])
                     .c_str());
  std::fprintf(f, "\n");  // the reference ends the file with an empty line
  std::fclose(f);
  double total = 0;
  for (double s : seconds) total += s;
  std::printf("%s: %d queries, %.6f s, %s\n", variant.c_str(), n, total, out.c_str());
  return 0;
}
