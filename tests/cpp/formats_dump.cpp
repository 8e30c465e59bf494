// Reads reference-format files with include/gst_formats.hpp and prints a digest
// that the Python readers (gpu4gst/formats.py) must reproduce:
//   formats_dump in  <file.in>   -> "graph V E sum_col sum_w rowsorted"
//   formats_dump bin <prefix>    -> "graph V E sum_col sum_w"
//   formats_dump g   <file.g>    -> "groups G nonempty members sum_members"
//   formats_dump q   <file.csv>  -> "queries n ids sum_ids"
#include <cstdio>
#include <string>

#include "gst_formats.hpp"

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  const std::string mode = argv[1], path = argv[2];
  std::string err;
  if (mode == "in" || mode == "bin") {
    gst_fmt::csr64 c;
    int rowsorted = 1;
    if (mode == "in") {
      gst_fmt::adjacency adj;
      if (!gst_fmt::read_in(path, adj, &err)) return std::fprintf(stderr, "%s\n", err.c_str()), 1;
      c = gst_fmt::csr_from_adjacency(adj);
      for (const auto& row : adj)
        for (size_t i = 1; i < row.size(); ++i)
          if (row[i] < row[i - 1]) rowsorted = 0;
    } else if (!gst_fmt::read_bin(path, c, &err)) {
      return std::fprintf(stderr, "%s\n", err.c_str()), 1;
    }
    long long sc = 0, sw = 0, sx = 0;
    for (int64_t i = 0; i < c.entries(); ++i) {
      sc += c.col[i];
      sw += c.w[i];
      sx += (long long)(i % 7 + 1) * c.col[i];  // order-sensitive
    }
    std::printf("graph %lld %lld %lld %lld %lld %d\n", (long long)c.vertices(),
                (long long)c.entries(), sc, sw, sx, rowsorted);
  } else if (mode == "g") {
    gst_fmt::adjacency g;
    if (!gst_fmt::read_groups(path, g, &err)) return std::fprintf(stderr, "%s\n", err.c_str()), 1;
    long long nonempty = 0, members = 0, sum = 0, sx = 0;
    for (size_t i = 0; i < g.size(); ++i) {
      nonempty += !g[i].empty();
      members += (long long)g[i].size();
      for (size_t j = 0; j < g[i].size(); ++j) {
        sum += g[i][j].first;
        sx += (long long)(j % 5 + 1) * g[i][j].first;
      }
    }
    std::printf("groups %zu %lld %lld %lld %lld\n", g.size(), nonempty, members, sum, sx);
  } else if (mode == "q") {
    std::vector<std::vector<int>> q;
    if (!gst_fmt::read_queries(path, q, &err)) return std::fprintf(stderr, "%s\n", err.c_str()), 1;
    long long ids = 0, sum = 0, sx = 0;
    for (const auto& one : q)
      for (size_t j = 0; j < one.size(); ++j) {
        ++ids;
        sum += one[j];
        sx += (long long)(j + 1) * one[j];
      }
    std::printf("queries %zu %lld %lld %lld\n", q.size(), ids, sum, sx);
  } else {
    return 2;
  }
  return 0;
}
