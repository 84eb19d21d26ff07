/* Host-only test program of the CPU oracle (oracle/oracle.c), built with
 * AddressSanitizer and UndefinedBehaviorSanitizer by
 * tests/test_host_sanitizers.py: the five enumerators on seeded random small
 * graphs (half of them with nested rows) must agree, with list output on,
 * through their worker threads. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  uint64_t maximal, non_maximal, hash_sum, hash_xor, nodes, b_alg, tasks, splits, skipped, cut;
  double seconds;
} orc_result;
typedef struct orc_list orc_list;
orc_list* orc_list_new(void);
void orc_list_free(orc_list* l);
int orc_gmbe(uint32_t, uint32_t, const int64_t*, const uint32_t*, const uint32_t*, const uint32_t*,
             int, int, int, int, int, int, uint32_t, uint32_t, double, orc_result*, orc_list*);
int orc_mbea(uint32_t, uint32_t, const int64_t*, const uint32_t*, const uint32_t*, const uint32_t*,
             int, int, uint32_t, uint32_t, double, orc_result*, orc_list*);
int orc_imbea(uint32_t, uint32_t, const int64_t*, const uint32_t*, const uint32_t*,
              const uint32_t*, int, int, uint32_t, uint32_t, double, orc_result*, orc_list*);
int orc_pmbe(uint32_t, uint32_t, const int64_t*, const uint32_t*, const uint32_t*,
             const uint32_t*, int, int, uint32_t, uint32_t, double, orc_result*, orc_list*);
int orc_brute(uint32_t, uint32_t, const int64_t*, const uint32_t*, const uint32_t*, const uint32_t*,
              int, orc_result*, orc_list*);

static uint64_t rng = 88172645463325252ull;
static uint32_t next_u32(void) {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return (uint32_t)rng;
}

int main(void) {
  int fails = 0;
  for (int t = 0; t < 60; ++t) {
    uint32_t nU = 2 + next_u32() % 40, nV = 2 + next_u32() % 20;
    uint32_t pct = 10 + next_u32() % 60;
    int64_t* off = malloc((nV + 1) * sizeof(int64_t));
    uint32_t* adj = malloc((size_t)nU * nV * sizeof(uint32_t) + 4);
    uint32_t* uo = malloc(nU * sizeof(uint32_t));
    uint32_t* vo = malloc(nV * sizeof(uint32_t));
    for (uint32_t u = 0; u < nU; ++u) uo[u] = u;
    for (uint32_t v = 0; v < nV; ++v) vo[v] = v;
    off[0] = 0;
    for (uint32_t v = 0; v < nV; ++v) {
      int64_t k = off[v];
      if (t % 2 && v > 0 && next_u32() % 10 < 7) {
        /* odd graphs: most rows are a subset of an earlier row (nested
         * neighbourhoods, which PMBE's containment order and skips act on) */
        uint32_t p = next_u32() % v;
        for (int64_t e = off[p]; e < off[p + 1]; ++e)
          if (next_u32() % 100 < 70) adj[k++] = adj[e];
      } else {
        for (uint32_t u = 0; u < nU; ++u)
          if (next_u32() % 100 < pct) adj[k++] = u;
      }
      off[v + 1] = k;
    }
    orc_result a, b, c, d, e, f;
    orc_list* la = orc_list_new();
    orc_list* lc = orc_list_new();
    orc_list* lf = orc_list_new();
    int rc = orc_gmbe(nU, nV, off, adj, uo, vo, 0, -1, 0, 1, 4, 2, 1, 0, 0.0, &a, la);
    rc |= orc_gmbe(nU, nV, off, adj, uo, vo, 0, 1, 2, 1, 4, 2, 1, 0, 0.0, &b, NULL);
    rc |= orc_mbea(nU, nV, off, adj, uo, vo, 0, 2, 1, 0, 0.0, &c, lc);
    rc |= orc_imbea(nU, nV, off, adj, uo, vo, 0, 2, 1, 0, 0.0, &d, NULL);
    rc |= orc_brute(nU, nV, off, adj, uo, vo, 0, &e, NULL);
    rc |= orc_pmbe(nU, nV, off, adj, uo, vo, 0, 2, 1, 0, 0.0, &f, lf);
    if (rc || a.maximal != e.maximal || b.maximal != e.maximal || c.maximal != e.maximal ||
        d.maximal != e.maximal || a.hash_sum != e.hash_sum || d.hash_sum != e.hash_sum ||
        c.hash_xor != e.hash_xor || f.maximal != e.maximal || f.hash_sum != e.hash_sum ||
        f.hash_xor != e.hash_xor) {
      printf("mismatch on graph %d (%u x %u): rc %d, %llu %llu %llu %llu %llu %llu\n", t, nU, nV,
             rc, (unsigned long long)a.maximal, (unsigned long long)b.maximal,
             (unsigned long long)c.maximal, (unsigned long long)d.maximal,
             (unsigned long long)e.maximal, (unsigned long long)f.maximal);
      ++fails;
    }
    orc_list_free(la);
    orc_list_free(lc);
    orc_list_free(lf);
    free(off);
    free(adj);
    free(uo);
    free(vo);
  }
  printf("done, %d failures\n", fails);
  return fails != 0;
}
