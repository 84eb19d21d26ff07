#include <cstdio>
#include <cstdlib>
#include <cstring>
// Host-only test program of the loader, writers and generators, built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_host_sanitizers.py.
#include "gmbe.h"
#include "../../gmbe_amd/csrc/graph.hpp"
#include <stdexcept>
#include <string>
static int fails = 0;
#define CHECK(x) do { if (!(x)) { printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #x); ++fails; } } while (0)
int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  const char* files[] = {"0 1 2\n\n3\n2 2 1\n", "", "\n\n\n", "5\n", "1 x 2\n", "4294967295\n", "0 1\n1 0\n"};
  for (int i = 0; i < 7; ++i) {
    const std::string ps = dir + "/f" + std::to_string(i) + ".adj";
    const char* path = ps.c_str();
    FILE* fp = fopen(path, "w"); fputs(files[i], fp); fclose(fp);
    for (int t = 0; t < 3; ++t) for (int o = 0; o < 4; ++o) for (int f = 0; f < 2; ++f) {
      gmbe_option opt = {t, o, f};
      gmbe_graph* g = nullptr;
      int rc = gmbe_graph_read_adj(path, &opt, &g);
      if (rc) { CHECK(g == nullptr); continue; }
      int64_t st[8]; CHECK(gmbe_graph_stats(g, st) == 0);
      gmbe_graph* d = nullptr;
      gmbe_option o2 = {1, 1, 1};
      if (gmbe_graph_derive(g, &o2, &d) == 0) {
        CHECK(gmbe_graph_write_edges(d, (dir + "/out.graph").c_str()) == 0);
        CHECK(gmbe_graph_write_adj(d, (dir + "/out.adj").c_str()) == 0);
        gmbe_graph_destroy(d);
      }
      gmbe_graph_destroy(g);
    }
  }
  for (uint64_t s = 0; s < 40; ++s) {
    gmbe_graph* g = nullptr;
    CHECK(gmbe_generate_small((4ull << 32) | s, 50, 2000, 3.0, 20.0, 2.2, &g) == 0);
    int64_t st[8]; gmbe_graph_stats(g, st);
    for (uint32_t u = 0; u < (uint32_t)st[0]; u += 97) {
      gmbe_graph* e = nullptr;
      if (gmbe_ego_subgraph(g, u, 3000, &e) == 0) gmbe_graph_destroy(e);
    }
    gmbe_graph_destroy(g);
  }
  gmbe_gen_params p = {};
  p.kind = 3; p.nU = 3000; p.nV = 20000; p.m = 60000; p.gamma_u = p.gamma_v = 2.0; p.cap = 2000;
  p.blocks = 20; p.block_min = 10; p.block_max = 40; p.drop = 0.05; p.seed = 3;
  gmbe_graph* g = nullptr;
  CHECK(gmbe_generate(&p, &g) == 0);
  gmbe_graph_destroy(g);
  // an exception inside an entry point's body becomes an error code and message
  CHECK(gmbe::guarded("probe_a", []() -> int { throw std::bad_alloc(); }) == GMBE_ERR_NOMEM);
  CHECK(strstr(gmbe_last_error(), "probe_a") != nullptr);
  CHECK(gmbe::guarded("probe_b", []() -> int { throw std::length_error("too long"); }) ==
        GMBE_ERR_CAPACITY);
  CHECK(gmbe::guarded("probe_c", []() -> int { throw std::runtime_error("boom"); }) ==
        GMBE_ERR_INTERNAL);
  CHECK(strstr(gmbe_last_error(), "boom") != nullptr);
  CHECK(gmbe::guarded("probe_d", []() -> int { throw 7; }) == GMBE_ERR_INTERNAL);
  CHECK(gmbe::guarded("probe_e", []() -> int { return 5; }) == 5);
  printf("done, %d failures\n", fails);
  return fails != 0;
}
