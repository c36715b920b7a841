// st-con: the ST-CON command line (bernaschi/.../st-conn/main.c:136-249) on
// fastcon, one GPU.
//
//   st-con -S SCALE [-E edgefactor] [-1 seed1] [-2 seed2] [-N pairs] [-G 0|1]
//          [-f edge_list_file] [-d stats_dir] [-R red -B blue] [-V num] [-v]
//          (-t -b -D -T -s -r -m are accepted for compatibility; they set
//           CUDA launch geometry and MPI buffers, which this engine does not
//           have)
//
// -V num (default 1, as the reference; 0 turns it off) validates up to num
// matching nodes per connected pair with validate_stcon_bfs's rule
// (st-conn/make_stcon.c:516-575,1578-1641): a BFS from the matching node,
// whose tree parents from t and from s must reach it within DEPTH steps,
// the longer walk taking exactly DEPTH. This engine reports one matching
// node per pair. A failed check prints the reference's ERROR line and stops
// the run with exit status 2.
//
// For each pair it prints the reference's START / ST-CON RUN / END lines and
// writes an st_stats_<date>_<time> file in the reference's format
// (st-conn/main.c:308-335, st-conn/make_stcon.c:414,495-500,623-658): one
// row per pair with DEPTH = the level at which the two frontiers meet
// (ceil(d/2) for distance d), MN-FOUND, the query time and TEPS, then the
// time, CONPS and visited-vertex statistics (sample standard deviation, as
// get_statistics, st-conn/reference_functions.c:531-557).
// Differences: -S graphs come from this engine's seeded R-MAT generator
// (initiator .57/.19/.19/.05, as the reference) or, with -G 1, its
// Erdős–Rényi generator, not from the Graph500 Kronecker generator, so the
// same seeds give a different graph; pairs follow find_stcon_roots's rule
// (degree > 0 and a non-self edge) from this engine's seeded stream; the
// default seeds are 1 and 2 instead of time-based ones. TEPS uses the
// visited vertices' degree sum halved (each undirected edge once).
#include <getopt.h>
#include <unistd.h>

#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "fastcon.h"

namespace {

void check(fc_status st, const char* what) {
  if (st != FC_OK) {
    std::fprintf(stderr, "%s: %s\n", what, fc_last_error());
    std::exit(1);
  }
}

struct Stats {
  double min, max, mean, std;
};

Stats statistics(const std::vector<double>& x) {
  Stats r{0, 0, 0, 0};
  if (x.empty()) return r;
  double m = 0;
  for (double v : x) m += v;
  m /= (double)x.size();
  double var = 0;
  for (double v : x) var += (v - m) * (v - m);
  r.mean = m;
  r.std = x.size() > 1 ? std::sqrt(var / (double)(x.size() - 1)) : 0.0;
  r.min = *std::min_element(x.begin(), x.end());
  r.max = *std::max_element(x.begin(), x.end());
  return r;
}

void usage(const char* prog) {
  std::fprintf(stderr,
               "Usage: %s -S SCALE -E Edgefactor -1 seed1 -2 seed2 -N num_pairs [-G 0|1] [-f filename]"
               " [-d stats_dir] [-R vertex -B vertex] [-V num] [-v]\n"
               "SCALE = log_2(# vertices) [integer, required unless -f]\n"
               "  edgefactor = (# edges) / (# vertices) = .5 * (average vertex degree) [float, defaults to 16]\n"
               "-f filename to run on the real graph stored in filename (edge list)\n"
               "-d directory to store the st_stats file\n"
               "-G (0|1) 0 for RMAT graph and 1 for Random graph\n"
               "-R vertex -B vertex (red and blue vertices)\n"
               "-V num (Maximum number of matching nodes to validate, default 1)\n",
               prog);
}

}  // namespace

int main(int argc, char* argv[]) {
  int scale = 0, num_pairs = 64, gen_type = 0, errflg = 0, validation = 1;
  double edgefactor = 16.0;
  uint64_t seed1 = 1, seed2 = 2;
  long long red = -1, blue = -1;
  std::string file, outdir = ".";
  int opt;
  while ((opt = getopt(argc, argv, "S:E:t:b:1:2:N:V:r:m:s:T:D:f:d:B:R:G:v")) != -1) {
    switch (opt) {
      case 'R': red = std::atoll(optarg); break;
      case 'B': blue = std::atoll(optarg); break;
      case 'S': scale = std::atoi(optarg); break;
      case 'E': edgefactor = std::atof(optarg); break;
      case '1': seed1 = std::strtoull(optarg, nullptr, 10); break;
      case '2': seed2 = std::strtoull(optarg, nullptr, 10); break;
      case 'N': num_pairs = std::atoi(optarg); break;
      case 'f': file = optarg; break;
      case 'd': outdir = optarg; break;
      case 'G': gen_type = std::atoi(optarg); break;
      case 'r':
      case 's':
      case 'm':
        if (std::atof(optarg) <= 0) {
          std::fprintf(stderr, "Option -%c requires a positive value\n", opt);
          errflg++;
        }
        break;
      case 'V': validation = std::atoi(optarg); break;
      case 't': case 'b': case 'D': case 'T': break;
      case 'v': std::printf("Version %s\n", fc_version()); return 0;
      default: errflg++; break;
    }
  }
  if ((file.empty() && scale <= 0) || edgefactor <= 0 || num_pairs <= 0 || errflg) {
    usage(argv[0]);
    return 1;
  }
  if (red != -1 && blue != -1) num_pairs = 1;

  fc_host_graph hg;
  if (!file.empty()) {
    int64_t* edges = nullptr;
    int64_t m = 0, nv = 0;
    check(fc_read_edge_list(file.c_str(), &edges, &m, &nv), "read edge list");
    check(fc_csr_from_edges(edges, m, nv, &hg), "build graph");
    fc_free(edges);
  } else if (gen_type == 0) {
    check(fc_gen_rmat(scale, edgefactor, 0.57, 0.19, 0.19, seed1 * 0x100000001B3ull ^ seed2, &hg), "generate R-MAT");
  } else {
    check(fc_gen_er(1LL << scale, 2.0 * edgefactor, seed1 * 0x100000001B3ull ^ seed2, &hg), "generate random graph");
  }

  char stamp[80];
  std::time_t now = std::time(nullptr);
  std::strftime(stamp, sizeof(stamp), "%Y%m%d_%H%M%S", std::localtime(&now));
  const std::string fname = outdir + "/st_stats_" + stamp;
  FILE* fs = std::fopen(fname.c_str(), "w");
  if (!fs) {
    std::fprintf(stderr, "Error opening file %s", fname.c_str());
    return 1;
  }
  std::fprintf(fs, "SCALE %d\n", file.empty() ? scale : (int)std::ceil(std::log2((double)std::max<int64_t>(hg.V, 1))));
  std::fprintf(fs, "Edgefactor %.1f\n", file.empty() ? edgefactor : (double)hg.E / 2.0 / (double)std::max<int64_t>(hg.V, 1));
  std::fprintf(fs, "MPI-Nodes %d\n", 1);
  std::fprintf(fs, "Num-ST-CON=%d\n", num_pairs);
  if (file.empty())
    std::fprintf(fs, "Graph-type %s\n", gen_type ? "Random" : "RMAT");
  else
    std::fprintf(fs, "Filename %s\n", file.c_str());
  std::fprintf(fs, "Seed1 %" PRIu64 "\n", seed1);
  std::fprintf(fs, "Seed2 %" PRIu64 "\n", seed2);

  std::vector<int32_t> s(num_pairs), t(num_pairs);
  if (red != -1 && blue != -1) {
    if (red < 0 || red >= hg.V || blue < 0 || blue >= hg.V) {
      std::fprintf(stderr, "red/blue vertex out of range [0,%" PRId64 ")\n", hg.V);
      return 1;
    }
    s[0] = (int32_t)red;
    t[0] = (int32_t)blue;
  } else {
    std::vector<int32_t> roots(2 * (size_t)num_pairs);
    check(fc_sample_roots(&hg, 2 * num_pairs, seed2, roots.data()), "sample pairs");
    for (int i = 0; i < num_pairs; ++i) {
      s[i] = roots[i];
      t[i] = roots[i + num_pairs];
    }
  }

  fc_graph* g = nullptr;
  if (hg.E < INT32_MAX) {
    std::vector<int32_t> off32((size_t)hg.V + 1);
    for (int64_t v = 0; v <= hg.V; ++v) off32[(size_t)v] = (int32_t)hg.offsets[v];
    check(fc_graph_create_csr32(off32.data(), hg.columns, (int32_t)hg.V, hg.E, 0, &g), "upload graph");
  } else {
    check(fc_graph_create_csr64(hg.offsets, hg.columns, hg.V, hg.E, 0, &g), "upload graph");
  }
  {  // first query of a process pays one-off costs; keep it out of the statistics
    int32_t c, d, mn;
    fc_stats st;
    check(fc_stcon(g, s[0], t[0], &c, &d, &mn, &st), "st-con");
  }

  std::fprintf(fs, "ST-CON-RUN RED-V BLUE-V DEPTH MN-FOUND TIME(Sec) TEPS\n");
  std::vector<double> times, conps, visited;
  int max_level = 0;
  for (int i = 0; i < num_pairs; ++i) {
    std::fprintf(stdout, "\nSTART %d ST-CON , RED = %d, BLUE = %d\n", i, s[i], t[i]);
    int32_t conn = 0, dist = -1, meet = -1;
    fc_stats st;
    check(fc_stcon(g, s[i], t[i], &conn, &dist, &meet, &st), "st-con");
    const double sec = st.time_ms / 1e3;
    const int depth = conn ? (dist + 1) / 2 : st.levels;
    const double teps = sec > 0 ? ((double)st.visited_edges / 2.0) / sec : 0.0;
    max_level = std::max(max_level, depth);
    if (!conn) {
      std::fprintf(stdout, "END %d ST-CON: NO MATCHING NODE FOUND\n", i);
    } else {
      std::fprintf(stdout, "ST-CON RUN %d: Depth = %d Global Matching Nodes = %d\n", i, depth, 1);
    }
    std::fprintf(fs, "%d %d %d %d %d %g %g\n", i, s[i], t[i], depth, conn ? 1 : 0, sec, teps);
    std::fprintf(stdout, "END STCON [%d]\n", i);
    if (conn && validation > 0 && s[i] != t[i]) {
      const fc_status vs = fc_validate_stcon(g, s[i], t[i], meet, depth);
      if (vs == FC_OK) {
        std::fprintf(stdout, "ST-CON Validation %d OK with Matching node %d\n", 0, meet);
      } else {
        std::fprintf(stderr, "%s\n", fc_last_error());
        std::fprintf(stdout, "ST-CON Validation %d ERROR with Matching node %d\n", 0, meet);
        std::fprintf(stdout, "seed1=%" PRIu64 ", seed2=%" PRIu64 "\n", seed1, seed2);
        std::fclose(fs);
        fc_graph_destroy(g);
        fc_host_graph_free(&hg);
        return 2;
      }
    }
    times.push_back(sec);
    conps.push_back(sec > 0 ? 1.0 / sec : 0.0);
    visited.push_back((double)st.visited_vertices);
  }

  const Stats tm = statistics(times), cp = statistics(conps), vs = statistics(visited);
  std::fprintf(stdout, "STCON STATS\n");
  std::fprintf(stdout, "N ST-CON:                       %d\n", num_pairs);
  std::fprintf(stdout, "max ST-CON level:               %d\n", max_level);
  std::fprintf(stdout, "min_time:                       %g s\n", tm.min);
  std::fprintf(stdout, "max_time:                       %g s\n", tm.max);
  std::fprintf(stdout, "mean_time:                      %g s\n", tm.mean);
  std::fprintf(stdout, "stddev_time:                    %g\n", tm.std);
  std::fprintf(stdout, "min_CONPS:                       %g \n", cp.min);
  std::fprintf(stdout, "max_CONPS:                       %g \n", cp.max);
  std::fprintf(stdout, "mean_CONPS:                      %g \n", cp.mean);
  std::fprintf(stdout, "stddev_CONPS:                    %g \n", cp.std);
  std::fprintf(stdout, "-----ST_CON_VISITED_NODES-----\n");
  std::fprintf(stdout, "min_visited:                       %g \n", vs.min);
  std::fprintf(stdout, "max_visited:                       %g \n", vs.max);
  std::fprintf(stdout, "mean_visited:                      %g \n", vs.mean);
  std::fprintf(stdout, "stddev_visited:                    %g \n", vs.std);

  std::fprintf(fs, "max-ST-CON-level %d\n", max_level);
  std::fprintf(fs, "min_time %g s\n", tm.min);
  std::fprintf(fs, "max_time %g s\n", tm.max);
  std::fprintf(fs, "mean_time %g s\n", tm.mean);
  std::fprintf(fs, "stddev_time %g\n", tm.std);
  std::fprintf(fs, "min_CONPS                   %g \n", cp.min);
  std::fprintf(fs, "max_CONPS                   %g \n", cp.max);
  std::fprintf(fs, "mean_CONPS                  %g \n", cp.mean);
  std::fprintf(fs, "stddev_CONPS                %g \n", cp.std);
  std::fprintf(fs, "mean_ST, mean_time, stddev_ST: %g, %g, %g \n", cp.mean, tm.mean, cp.std);
  std::fprintf(fs, "-----ST_CON_VISITED_NODES-----\n");
  std::fprintf(fs, "min_visited, max_visited, mean_visited, stddev_visited: %g, %g, %g, %g\n", vs.min, vs.max, vs.mean,
               vs.std);
  std::fclose(fs);
  std::printf("stats written to %s\n", fname.c_str());
  fc_graph_destroy(g);
  fc_host_graph_free(&hg);
  return 0;
}
