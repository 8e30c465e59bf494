// gst_formats.hpp - readers and writers for the reference's on-disk formats.
//
// Header-only C++17, no device code. Each function returns false and sets
// *err on a malformed or missing file (the reference readers call exit()).
//
//   <name>_beg_pos.bin / _csr.bin / _weight.bin
//       little-endian int64 arrays beg_pos[V+1], csr[E], weight[E] with
//       E = beg_pos[V]; V is file size / 8 - 1 (code/TrimCDP-WB/include/graph.hpp:25-113,
//       writer data/tuple_text_to_bin.cpp:205-237, both directions of every edge).
//   <name>.in
//       first line starts with V; every later line with exactly three integers
//       "v1 v2 w" adds both directions; each row is then sorted by (neighbour,
//       weight) (read_input_graph, code/TrimCDP/include/
//       graph_hash_of_mixed_weighted_read_for_GSTP.h:75-148).
//   <name>.g
//       one group per line, "gN:v v v ..." (a trailing space is allowed); line
//       gN is group id N-1; members keep their order, sorted by vertex id
//       (read_Group, same file:190-239; code/TrimCDP-WB/include/read_group.hpp:9-53).
//       Unlike the reference reader, ids may come in any order and gaps are
//       empty groups.
//   <name><k>.csv
//       one query per line, space-separated 0-based group ids
//       (read_inquire, same file:240-286; read_group.hpp:55-89).
//   result CSV
//       the per-variant header lines and row layout of the reference harnesses
//       (code/TrimCDP/include/exp_GPU_nonHop.h:83-85, 129;
//       code/D-TrimCDP/include/exp_GPU_Hop.h:80-82, 112;
//       code/TrimCDP-WB/src/GSTnonHop.cu:241-243, 493;
//       code/D-TrimCDP-WB/src/GPUHop.cu:157-159, 262).
#ifndef GST_FORMATS_HPP
#define GST_FORMATS_HPP

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

namespace gst_fmt {

typedef std::vector<std::vector<std::pair<int, int>>> adjacency;

struct csr64 {
  std::vector<int64_t> beg, col, w;
  int64_t vertices() const { return beg.empty() ? 0 : (int64_t)beg.size() - 1; }
  int64_t entries() const { return beg.empty() ? 0 : beg.back(); }
};

inline bool fail(std::string* err, const std::string& msg) {
  if (err) *err = msg;
  return false;
}

inline bool read_i64_file(const std::string& path, std::vector<int64_t>& out, std::string* err) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return fail(err, "cannot open " + path);
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz < 0 || sz % 8) {
    std::fclose(f);
    return fail(err, path + ": size is not a multiple of 8 bytes");
  }
  out.resize((size_t)sz / 8);
  const size_t got = out.empty() ? 0 : std::fread(out.data(), 8, out.size(), f);
  std::fclose(f);
  if (got != out.size()) return fail(err, path + ": short read");
  return true;
}

inline bool write_i64_file(const std::string& path, const std::vector<int64_t>& a,
                           std::string* err) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return fail(err, "cannot create " + path);
  const size_t put = a.empty() ? 0 : std::fwrite(a.data(), 8, a.size(), f);
  const bool ok = std::fclose(f) == 0 && put == a.size();
  return ok ? true : fail(err, path + ": short write");
}

// <prefix>_beg_pos.bin, <prefix>_csr.bin, <prefix>_weight.bin
inline bool read_bin(const std::string& prefix, csr64& g, std::string* err) {
  if (!read_i64_file(prefix + "_beg_pos.bin", g.beg, err)) return false;
  if (g.beg.size() < 2) return fail(err, prefix + "_beg_pos.bin: needs at least 2 entries");
  const int64_t E = g.beg.back();
  if (g.beg[0] != 0 || E <= 0) return fail(err, prefix + "_beg_pos.bin: bad first/last offset");
  for (size_t i = 1; i < g.beg.size(); ++i)
    if (g.beg[i] < g.beg[i - 1]) return fail(err, prefix + "_beg_pos.bin: offsets decrease");
  if (!read_i64_file(prefix + "_csr.bin", g.col, err)) return false;
  if (!read_i64_file(prefix + "_weight.bin", g.w, err)) return false;
  // the reference reads exactly beg_pos[V] entries of each (graph.hpp:62-68, 85-91)
  if ((int64_t)g.col.size() < E || (int64_t)g.w.size() < E)
    return fail(err, prefix + ": csr/weight files hold fewer than beg_pos[V] entries");
  g.col.resize(E);
  g.w.resize(E);
  const int64_t V = g.vertices();
  for (int64_t i = 0; i < E; ++i)
    if (g.col[i] < 0 || g.col[i] >= V) return fail(err, prefix + "_csr.bin: vertex id out of range");
  return true;
}

inline bool write_bin(const std::string& prefix, const csr64& g, std::string* err) {
  return write_i64_file(prefix + "_beg_pos.bin", g.beg, err) &&
         write_i64_file(prefix + "_csr.bin", g.col, err) &&
         write_i64_file(prefix + "_weight.bin", g.w, err);
}

// splits on any run of the separator characters
inline std::vector<std::string> split(const std::string& s, const char* seps) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < s.size()) {
    while (i < s.size() && std::strchr(seps, s[i])) ++i;
    size_t j = i;
    while (j < s.size() && !std::strchr(seps, s[j])) ++j;
    if (j > i) out.push_back(s.substr(i, j - i));
    i = j;
  }
  return out;
}

inline bool to_int(const std::string& s, long long& v) {
  if (s.empty()) return false;
  char* end = nullptr;
  errno = 0;
  v = std::strtoll(s.c_str(), &end, 10);
  return errno == 0 && end && *end == '\0';
}

inline bool read_in(const std::string& path, adjacency& g, std::string* err) {
  std::ifstream f(path);
  if (!f) return fail(err, "cannot open " + path);
  std::string line;
  if (!std::getline(f, line)) return fail(err, path + ": empty file");
  const auto head = split(line, " \t\r");
  long long V = 0;
  if (head.empty()) return fail(err, path + ": first line has no vertex count");
  // the reference parses the count with stod (read_for_GSTP.h:86)
  V = (long long)std::strtod(head[0].c_str(), nullptr);
  if (V <= 0 || V > (1LL << 31) - 1) return fail(err, path + ": bad vertex count");
  g.assign((size_t)V, {});
  long long lineno = 1;
  while (std::getline(f, line)) {
    ++lineno;
    const auto t = split(line, " \t\r");
    if (t.size() != 3) continue;  // the reference skips every other line shape
    long long a, b, w;
    if (!to_int(t[0], a) || !to_int(t[1], b) || !to_int(t[2], w))
      return fail(err, path + ": line " + std::to_string(lineno) + " is not three integers");
    if (a < 0 || b < 0 || a >= V || b >= V)
      return fail(err, path + ": line " + std::to_string(lineno) + " names a vertex out of range");
    g[a].push_back({(int)b, (int)w});
    g[b].push_back({(int)a, (int)w});
  }
  for (auto& row : g) std::sort(row.begin(), row.end());
  return true;
}

inline bool write_in(const std::string& path, int64_t V, const std::vector<int32_t>& src,
                     const std::vector<int32_t>& dst, const std::vector<int32_t>& w,
                     std::string* err) {
  std::ofstream f(path);
  if (!f) return fail(err, "cannot create " + path);
  f << V << " " << src.size() << "\n";
  for (size_t i = 0; i < src.size(); ++i) f << src[i] << " " << dst[i] << " " << w[i] << "\n";
  return f.good() ? true : fail(err, path + ": write failed");
}

inline bool read_groups(const std::string& path, adjacency& groups, std::string* err) {
  std::ifstream f(path);
  if (!f) return fail(err, "cannot open " + path);
  groups.clear();
  std::string line;
  long long lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.find_first_not_of(" \t") == std::string::npos) continue;
    const size_t colon = line.find(':');
    if (colon == std::string::npos || colon < 2 || line[0] != 'g')
      return fail(err, path + ": line " + std::to_string(lineno) + " is not gN:members");
    long long id;
    if (!to_int(line.substr(1, colon - 1), id) || id < 1)
      return fail(err, path + ": line " + std::to_string(lineno) + " has a bad group number");
    if ((size_t)id > groups.size()) groups.resize((size_t)id);
    auto& g = groups[(size_t)id - 1];
    g.clear();
    for (const auto& s : split(line.substr(colon + 1), " \t")) {
      long long v;
      if (!to_int(s, v) || v < 0)
        return fail(err, path + ": line " + std::to_string(lineno) + " has a bad member");
      g.push_back({(int)v, 1});
    }
    std::stable_sort(g.begin(), g.end(), [](const std::pair<int, int>& a,
                                            const std::pair<int, int>& b) {
      return a.first < b.first;
    });
  }
  return true;
}

inline bool read_queries(const std::string& path, std::vector<std::vector<int>>& q,
                         std::string* err) {
  std::ifstream f(path);
  if (!f) return fail(err, "cannot open " + path);
  q.clear();
  std::string line;
  long long lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    const auto t = split(line, " \t\r,");
    if (t.empty()) continue;
    std::vector<int> one;
    for (const auto& s : t) {
      long long v;
      if (!to_int(s, v) || v < 0)
        return fail(err, path + ": line " + std::to_string(lineno) + " has a bad group id");
      one.push_back((int)v);
    }
    q.push_back(one);
  }
  return true;
}

// symmetric CSR (int64) from an adjacency whose rows already hold both
// directions, in row order
inline csr64 csr_from_adjacency(const adjacency& g) {
  csr64 c;
  c.beg.assign(g.size() + 1, 0);
  for (size_t v = 0; v < g.size(); ++v) {
    c.beg[v + 1] = c.beg[v] + (int64_t)g[v].size();
    for (const auto& e : g[v]) {
      c.col.push_back(e.first);
      c.w.push_back(e.second);
    }
  }
  return c;
}

inline void groups_csr(const adjacency& groups, std::vector<int64_t>& off,
                       std::vector<int32_t>& mem) {
  off.assign(groups.size() + 1, 0);
  mem.clear();
  for (size_t i = 0; i < groups.size(); ++i) {
    off[i + 1] = off[i] + (int64_t)groups[i].size();
    for (const auto& m : groups[i]) mem.push_back(m.first);
  }
}

// Result CSV of the four reference harnesses. variant: "TrimCDP", "D-TrimCDP",
// "TrimCDP-WB" or "D-TrimCDP-WB".
inline std::string result_file_name(const std::string& variant, const std::string& data, int T,
                                    int D, int first, int last) {
  std::string s = variant + "_" + data + "_T" + std::to_string(T) + "_";
  if (variant.rfind("D-", 0) == 0) s += "D" + std::to_string(D) + "_";
  return s + std::to_string(first) + "-" + std::to_string(last) + ".csv";
}

inline std::string result_header(const std::string& variant) {
  if (variant == "TrimCDP")
    return "task_ID,task,GPU1_nonHop_time,GPU_nonHop_cost,GPU_nonHop_memory,counts,process_num,"
           "mid_counts,mid_process_num,mid_time";
  if (variant == "D-TrimCDP")
    return "task_ID,task,GPU1_Hop_time,GPU1_Hop_cost,GPU1_Hop_memory,counts,process_num,"
           "mid_counts,mid_process_num,mid_time";
  if (variant == "TrimCDP-WB")
    return "task_ID,task,GPU2_nonHop_time,GPU2_nonHop_cost,GPU2_nonHop_memory,counts,process_num";
  return "task_ID,task,GPU2_Hop_time,GPU2_Hop_cost,GPU2_Hop_memory,counts,process_num";
}

// one row: the task column is the query's group ids, each followed by a space
// (exp_GPU_nonHop.h:113); time with 8 fixed decimals (exp_GPU_nonHop.h:78-80)
inline std::string result_row(const std::string& variant, int task_id,
                              const std::vector<int>& query, double seconds, long long cost,
                              long long memory, long long counts, long long processed) {
  std::string task;
  for (int g : query) task += std::to_string(g) + " ";
  char t[64];
  std::snprintf(t, sizeof t, "%.8f", seconds);
  std::string s = std::to_string(task_id) + "," + task + "," + t + "," + std::to_string(cost) +
                  "," + std::to_string(memory) + "," + std::to_string(counts) + "," +
                  std::to_string(processed);
  if (variant == "TrimCDP" || variant == "D-TrimCDP") s += ",0,0,0.00000000";
  return s;
}

}  // namespace gst_fmt

#endif
