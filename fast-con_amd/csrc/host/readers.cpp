// Graph file readers for every format of BFS-4K/FileFormat.txt, with the
// reference's vertex numbering, direction rules and adjacency order:
//   format sniff + header   BFS-4K/Util/readGraphHeader.cpp:14-114
//   bodies                  BFS-4K/Util/readGraphSTD.cpp:14-145
//   SNAP first-seen remap   BFS-4K/Util/readUtil.cpp:3-11
//   COO -> CSR              BFS-4K/src/host/graph.cpp:19-42 (counting sort,
//                           adjacency in file order, both directions when
//                           undirected, no self-loop or duplicate removal)
// plus the ST-CON edge-list loader (st-conn/load_graph.c:20-104).
#include "common.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

using fc::fail;

namespace fc {
fc_status alloc_host_graph(int64_t V, int64_t E, fc_host_graph* out);
}

namespace {

enum Dir { DIRECTED = 0, UNDIRECTED = 1, UNDEFINED = 2 };

void skip_line(std::ifstream& f) { f.ignore(std::numeric_limits<std::streamsize>::max(), '\n'); }

bool all_digits(const std::string& s) {
  return std::all_of(s.begin(), s.end(), [](unsigned char c) { return std::isdigit(c) != 0; });
}

// COO -> CSR exactly as Graph::ToCSR: degree count, prefix sum, then a fill
// in COO order; when undirected the reverse entry is placed right after.
fc_status to_csr(int64_t V, const std::vector<std::pair<int32_t, int32_t>>& coo, bool undirected,
                 fc_host_graph* out) {
  std::vector<int64_t> deg((size_t)V, 0);
  for (auto& e : coo) {
    if (e.first < 0 || e.first >= V || e.second < 0 || e.second >= V)
      return fail(FC_ERR_FORMAT, "vertex id out of range [0,%lld): edge (%d,%d)", (long long)V,
                  e.first, e.second);
    deg[e.first]++;
    if (undirected) deg[e.second]++;
  }
  int64_t E = 0;
  for (int64_t v = 0; v < V; ++v) E += deg[v];
  fc_status st = fc::alloc_host_graph(V, E, out);
  if (st != FC_OK) return st;
  out->offsets[0] = 0;
  for (int64_t v = 0; v < V; ++v) out->offsets[v + 1] = out->offsets[v] + deg[v];
  std::vector<int64_t> tmp((size_t)V, 0);
  for (auto& e : coo) {
    out->columns[out->offsets[e.first] + tmp[e.first]++] = e.second;
    if (undirected) out->columns[out->offsets[e.second] + tmp[e.second]++] = e.first;
  }
  out->directed = undirected ? 0 : 1;
  return FC_OK;
}

}  // namespace

extern "C" fc_status fc_read_graph(const char* path, fc_direction user_dir, fc_host_graph* out) {
  fc::clear_error();
  if (!path || !out) return fail(FC_ERR_INVALID, "fc_read_graph: null argument");
  std::memset(out, 0, sizeof(*out));
  struct stat sb;
  std::ifstream fin(path);
  if (!fin.is_open() || stat(path, &sb) != 0 || (sb.st_mode & S_IFMT) != S_IFREG)
    return fail(FC_ERR_IO, " Error. Read file: %s", path);

  std::string s;
  fin >> s;
  fin.clear();
  fin.seekg(0, std::ios::beg);

  enum Fmt { MM, D9, D10, SNAP, PGS } fmt;
  int64_t V = 0, nof_lines = 0;
  Dir file_dir = UNDEFINED;
  if (s == "c" || s == "p") {
    fmt = D9;
    while (fin.peek() == 'c') skip_line(fin);
    std::string nil;
    fin >> nil >> nil >> V >> nof_lines;
  } else if (s == "##") {
    fmt = PGS;
    skip_line(fin);
    file_dir = DIRECTED;
    int64_t p0 = 0, p1 = 0;
    fin >> p0 >> p1 >> nof_lines;
    V = p0 + p1;
  } else if (s == "%%MatrixMarket") {
    fmt = MM;
    std::string line;
    std::getline(fin, line);
    file_dir = line.find("symmetric") != std::string::npos ? UNDIRECTED : DIRECTED;
    while (fin.peek() == '%') skip_line(fin);
    std::string cols;
    fin >> V >> cols >> nof_lines;
  } else if (s == "#") {
    fmt = SNAP;
    std::string tmp;
    fin >> tmp >> tmp;
    file_dir = tmp == "Undirected" ? UNDIRECTED : DIRECTED;
    skip_line(fin);
    while (fin.peek() == '#') {
      std::getline(fin, tmp);
      if (tmp.size() >= 8 && tmp.substr(2, 6) == "Nodes:") {
        std::istringstream ss(tmp);
        ss >> tmp >> tmp >> V >> tmp >> nof_lines;
        break;
      }
    }
  } else if (s == "%" || all_digits(s)) {
    fmt = D10;
    while (fin.peek() == '%') skip_line(fin);
    std::string str;
    // The third token is read even when the header line has only two, as the
    // reference's `fin >> V >> nof_lines >> str` does.
    fin >> V >> nof_lines >> str;
    file_dir = str == "100" ? DIRECTED : UNDIRECTED;
  } else {
    return fail(FC_ERR_FORMAT, " Error. Graph Type not recognized: %s %s", path, s.c_str());
  }
  if (!fin || V <= 0 || V > INT32_MAX || nof_lines < 0)
    return fail(FC_ERR_FORMAT, "malformed header in %s", path);

  Dir dir;
  if (user_dir == FC_DIR_UNDIRECTED)
    dir = UNDIRECTED;
  else if (user_dir == FC_DIR_DIRECTED)
    dir = DIRECTED;
  else
    dir = file_dir == UNDEFINED ? DIRECTED : file_dir;

  fin.clear();
  fin.seekg(0, std::ios::beg);
  std::vector<std::pair<int32_t, int32_t>> coo;
  coo.reserve((size_t)std::min<int64_t>(nof_lines, 1ll << 28));

  switch (fmt) {
    case MM: {
      while (fin.peek() == '%') skip_line(fin);
      skip_line(fin);
      for (int64_t l = 0; l < nof_lines; ++l) {
        int64_t a, b;
        if (!(fin >> a >> b)) return fail(FC_ERR_FORMAT, "%s: truncated at edge %lld", path, (long long)l);
        coo.emplace_back((int32_t)(a - 1), (int32_t)(b - 1));
        skip_line(fin);
      }
      break;
    }
    case D9: {
      int c;
      std::string nil;
      while ((c = fin.peek()) != EOF) {
        if (c == 'a') {
          int64_t a, b;
          fin >> nil >> a >> b;
          coo.emplace_back((int32_t)(a - 1), (int32_t)(b - 1));
        }
        skip_line(fin);
      }
      break;
    }
    case D10: {
      // Each of the V body lines lists the 1-based neighbours of one vertex.
      // The reference builds the CSR as directed from this listing and marks
      // the graph undirected afterwards (readGraphSTD.cpp:36-48), so the
      // adjacency is exactly the listing.
      while (fin.peek() == '%') skip_line(fin);
      skip_line(fin);
      for (int64_t v = 0; v < V; ++v) {
        std::string line;
        if (!std::getline(fin, line)) break;
        std::istringstream ss(line);
        int64_t x;
        while (ss >> x) coo.emplace_back((int32_t)v, (int32_t)(x - 1));
      }
      fc_status st = to_csr(V, coo, false, out);
      if (st == FC_OK) out->directed = dir == DIRECTED ? 1 : 0;
      return st;
    }
    case SNAP: {
      while (fin.peek() == '#') skip_line(fin);
      std::unordered_map<int64_t, int32_t> map;
      auto id = [&](int64_t raw) {
        auto it = map.find(raw);
        if (it != map.end()) return it->second;
        int32_t k = (int32_t)map.size();
        map.emplace(raw, k);
        return k;
      };
      for (int64_t l = 0; l < nof_lines; ++l) {
        int64_t a, b;
        if (!(fin >> a >> b)) return fail(FC_ERR_FORMAT, "%s: truncated at edge %lld", path, (long long)l);
        int32_t ia = id(a);
        int32_t ib = id(b);
        coo.emplace_back(ia, ib);
      }
      break;
    }
    case PGS:
      return fail(FC_ERR_FORMAT, "%s: PGSolver bodies are not read (header only, as in the reference)", path);
  }
  return to_csr(V, coo, dir == UNDIRECTED, out);
}

extern "C" fc_status fc_read_edge_list(const char* path, int64_t** edges_out, int64_t* nedges_out,
                                       int64_t* nverts_out) {
  fc::clear_error();
  if (!path || !edges_out || !nedges_out || !nverts_out)
    return fail(FC_ERR_INVALID, "fc_read_edge_list: null argument");
  FILE* fp = std::fopen(path, "rt");
  if (!fp) return fail(FC_ERR_IO, "Error opening file %s", path);
  char line[256], token[64];
  unsigned long nodes = 0, nedges = 0;
  bool body = false;
  while (std::fgets(line, sizeof(line), fp)) {
    token[0] = 0;
    if (std::sscanf(line, "# %63s", token) == 1) {
      if (std::strcmp(token, "FromNodeId") == 0) {
        body = true;
        break;
      }
      if (std::strcmp(token, "Nodes:") == 0) std::sscanf(line, "# Nodes: %lu Edges: %lu ", &nodes, &nedges);
    }
  }
  if (!body) {
    std::fclose(fp);
    return fail(FC_ERR_FORMAT, "%s: no '# FromNodeId' header line", path);
  }
  std::vector<int64_t> e;
  e.reserve(2 * (size_t)nedges);
  while (std::fgets(line, sizeof(line), fp)) {
    long long u, v;
    if (std::sscanf(line, "%lld %lld", &u, &v) == 2) {
      e.push_back(u);
      e.push_back(v);
    }
  }
  std::fclose(fp);
  int64_t* buf = (int64_t*)std::malloc(sizeof(int64_t) * (e.size() > 0 ? e.size() : 1));
  if (!buf) return fail(FC_ERR_NOMEM, "host allocation failed");
  std::memcpy(buf, e.data(), sizeof(int64_t) * e.size());
  *edges_out = buf;
  *nedges_out = (int64_t)(e.size() / 2);
  *nverts_out = (int64_t)nodes;
  return FC_OK;
}
