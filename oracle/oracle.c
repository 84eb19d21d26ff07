/*
 * oracle.c — CPU restatement of GMBE's maximal biclique enumeration.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests, the smoke check
 * and the CPU-baseline leg of the benchmark call it; the engine never does.
 *
 * Independent enumerators over the same CSR input:
 *
 *   orc_gmbe   restates the reference's GPU traversal in its own order:
 *              task roots (IterFinderWithMultipleVertex,
 *              src/IterFinderGpu.cu:772-887), the stack machine (IterProcess,
 *              src/IterFinderGpu.cu:12-69) with finder_push / finder_pop and
 *              the PRUNE_EN rule (src/IterFinderGpu.h:206-303), the
 *              bound_height / bound_size split rule (src/IterFinderGpu.cu:835-858)
 *              with children built as in LargeToTiny (src/IterFinderGpu.cu:889-910),
 *              top-level vertices in descending id (src/IterFinderGpu.cu:1008).
 *              It also counts the algorithmic bytes B_alg of the unsplit
 *              traversal (one unit per task root and per finder_push).
 *   orc_mbea   restates the reference's CPU finder MbeaFinder
 *              (src/BicliqueFinder.cu:18-57) with the 2-hop top level of
 *              MbeaAdvFinder (baselines/MBE.zip:MBE/BicliqueFinder.cpp:42-88),
 *              multi-threaded over top-level vertices.
 *   orc_imbea  restates the authors' improved CPU finder ImbeaAdvFinder
 *              (baselines/MBE.zip:MBE/BicliqueFinder.cpp:227-350): candidates
 *              ordered by overlap, dominated candidates moved to the excluded
 *              set with the pushed vertex (and out of P by the reference's
 *              own merge walk); a further CPU baseline.
 *   orc_pmbe   restates the authors' CPU finder PmbeAdvFinder
 *              (baselines/MBE.zip:MBE/PMBE.cpp:3-77, 251-354): V renumbered
 *              by the neighbourhood-containment order, candidates skipped
 *              inside the range a done candidate released, maximality by
 *              N(L') \ R within the candidates; a further CPU baseline.
 *   orc_brute  checks every subset R of V (|V| <= 24, |U| <= 64): the
 *              definition of a maximal biclique, L = N(R) != {} and
 *              R = N(L).
 *
 * Counter representation: the reference packs level:16|count:16 in one int
 * (src/IterFinderGpu.h:195-203), which caps |L| at 65,535; here level and
 * count are separate int32 arrays with the same meaning (ACTIVE = 0x7fff and
 * TOUCHED = 0x3fff become INT32 sentinels).
 *
 * Hash of a biclique (frozen; the engine implements the same definition):
 *   key(side, x) = mix64((x << 1) | side), side 0 = original U, 1 = original V
 *   h = mix64(sum_A key(0,a) ^ rotl64(sum_B key(1,b), 31))
 * with mix64 the splitmix64 output function.  Digest = (count, sum h, xor h).
 *
 * Sampling and time limits: every enumerator takes the task roots of
 * processing ranks phase, phase + stride, ... (phase may exceed the stride).
 * time_limit stops claiming roots once passed; orc_mbea, orc_imbea and
 * orc_pmbe also abandon the subtrees they are in, so their counts are then
 * partial but their rate is that of the work done (the benchmark's CPU
 * baseline reads it that way), and orc_result.cut counts the threads so cut:
 * a sample is complete only when every root was claimed and cut is 0.
 *
 * Parity pinning of each function: see tests/test_oracle.py (closed forms,
 * brute force, the three-way agreement, and the counts of the reference's
 * own CPU code built from baselines/MBE.zip into oracle/_ref/).
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define LV_ACTIVE 0x7fffffff
#define LV_TOUCHED 0x3fffffff

typedef struct {
  uint64_t maximal, non_maximal, hash_sum, hash_xor, nodes, b_alg, tasks;
  uint64_t splits;  /* orc_gmbe: tasks split by the bound rule */
  uint64_t skipped; /* orc_pmbe: candidates skipped inside a containment range */
  uint64_t cut;     /* threads whose search the time limit cut (counts partial) */
  double seconds;
} orc_result;

/* growable u32 list of biclique records [nA, nB, A..., B...] (original ids) */
typedef struct {
  uint32_t* d;
  size_t n, cap;
} orc_list;

orc_list* orc_list_new(void) { return (orc_list*)calloc(1, sizeof(orc_list)); }
void orc_list_free(orc_list* l) {
  if (l) {
    free(l->d);
    free(l);
  }
}
size_t orc_list_size(const orc_list* l) { return l ? l->n : 0; }
const uint32_t* orc_list_data(const orc_list* l) { return l ? l->d : NULL; }

static void list_push(orc_list* l, uint32_t x) {
  if (l->n == l->cap) {
    l->cap = l->cap ? l->cap * 2 : 1024;
    l->d = (uint32_t*)realloc(l->d, l->cap * sizeof(uint32_t));
    if (!l->d) {
      fprintf(stderr, "oracle: out of memory\n");
      abort();
    }
  }
  l->d[l->n++] = x;
}
static void list_append(orc_list* dst, const orc_list* src) {
  for (size_t i = 0; i < src->n; ++i) list_push(dst, src->d[i]);
}

static inline uint64_t mix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

uint64_t orc_hash_pair(const uint32_t* A, size_t nA, const uint32_t* B, size_t nB) {
  uint64_t sa = 0, sb = 0;
  for (size_t i = 0; i < nA; ++i) sa += mix64(((uint64_t)A[i] << 1) | 0u);
  for (size_t i = 0; i < nB; ++i) sb += mix64(((uint64_t)B[i] << 1) | 1u);
  return mix64(sa ^ rotl64(sb, 31));
}

static double wall(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/* ------------------------------------------------------------------------ */
/* graph: V->U rows as given, U->V built here; per-vertex keys               */

typedef struct {
  uint32_t nU, nV;
  const int64_t* v_off;
  const uint32_t* v_adj;
  int64_t* u_off;
  uint32_t* u_adj;
  uint64_t* kU; /* key of internal U vertex (original side folded in) */
  uint64_t* kV;
  const uint32_t* u_orig;
  const uint32_t* v_orig;
  int swapped;
  uint32_t dmaxV; /* max degree of a V vertex */
} og;

static int og_init(og* g, uint32_t nU, uint32_t nV, const int64_t* v_off, const uint32_t* v_adj,
                   const uint32_t* u_orig, const uint32_t* v_orig, int swapped) {
  memset(g, 0, sizeof(*g));
  g->nU = nU;
  g->nV = nV;
  g->v_off = v_off;
  g->v_adj = v_adj;
  g->u_orig = u_orig;
  g->v_orig = v_orig;
  g->swapped = swapped;
  int64_t E = v_off[nV];
  g->u_off = (int64_t*)calloc((size_t)nU + 1, sizeof(int64_t));
  g->u_adj = (uint32_t*)malloc((size_t)(E ? E : 1) * sizeof(uint32_t));
  g->kU = (uint64_t*)malloc(((size_t)nU + 1) * sizeof(uint64_t));
  g->kV = (uint64_t*)malloc(((size_t)nV + 1) * sizeof(uint64_t));
  if (!g->u_off || !g->u_adj || !g->kU || !g->kV) return -1;
  for (uint32_t v = 0; v < nV; ++v) {
    if (v_off[v + 1] - v_off[v] > g->dmaxV) g->dmaxV = (uint32_t)(v_off[v + 1] - v_off[v]);
    for (int64_t e = v_off[v]; e < v_off[v + 1]; ++e) {
      if (v_adj[e] >= nU) return -2;
      if (e > v_off[v] && v_adj[e] <= v_adj[e - 1]) return -3; /* rows must be ascending */
      g->u_off[v_adj[e] + 1]++;
    }
  }
  for (uint32_t u = 0; u < nU; ++u) g->u_off[u + 1] += g->u_off[u];
  int64_t* pos = (int64_t*)malloc(((size_t)nU + 1) * sizeof(int64_t));
  memcpy(pos, g->u_off, ((size_t)nU + 1) * sizeof(int64_t));
  for (uint32_t v = 0; v < nV; ++v)
    for (int64_t e = v_off[v]; e < v_off[v + 1]; ++e) g->u_adj[pos[v_adj[e]]++] = v;
  free(pos);
  int su = swapped ? 1 : 0, sv = swapped ? 0 : 1;
  for (uint32_t u = 0; u < nU; ++u) g->kU[u] = mix64(((uint64_t)(u_orig ? u_orig[u] : u) << 1) | (uint64_t)su);
  for (uint32_t v = 0; v < nV; ++v) g->kV[v] = mix64(((uint64_t)(v_orig ? v_orig[v] : v) << 1) | (uint64_t)sv);
  return 0;
}
static void og_free(og* g) {
  free(g->u_off);
  free(g->u_adj);
  free(g->kU);
  free(g->kV);
}
static inline uint64_t hash_internal(const og* g, uint64_t sumL, uint64_t sumR) {
  uint64_t sa = g->swapped ? sumR : sumL, sb = g->swapped ? sumL : sumR;
  return mix64(sa ^ rotl64(sb, 31));
}
/* record internal (L, R) as original-id record [nA, nB, A, B] */
static void record(const og* g, orc_list* l, const uint32_t* L, size_t nL, const uint32_t* R,
                   size_t nR) {
  if (!l) return;
  const uint32_t *A = L, *B = R;
  size_t nA = nL, nB = nR;
  const uint32_t *mA = g->u_orig, *mB = g->v_orig;
  if (g->swapped) {
    A = R;
    nA = nR;
    B = L;
    nB = nL;
    mA = g->v_orig;
    mB = g->u_orig;
  }
  list_push(l, (uint32_t)nA);
  list_push(l, (uint32_t)nB);
  for (size_t i = 0; i < nA; ++i) list_push(l, mA ? mA[A[i]] : A[i]);
  for (size_t i = 0; i < nB; ++i) list_push(l, mB ? mB[B[i]] : B[i]);
}
static inline void add_biclique(orc_result* r, uint64_t h) {
  r->maximal++;
  r->hash_sum += h;
  r->hash_xor ^= h;
}

static inline int64_t bsearch_u32(const uint32_t* a, int64_t n, uint32_t key) {
  int64_t lo = 0, hi = n - 1;
  while (lo <= hi) {
    int64_t mid = (lo + hi) >> 1;
    uint32_t x = a[mid];
    if (x == key) return mid;
    if (x < key)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  return -1;
}
/* sorted intersection, out may alias neither input */
static int64_t intersect(const uint32_t* a, int64_t na, const uint32_t* b, int64_t nb, uint32_t* out) {
  int64_t i = 0, j = 0, n = 0;
  if (na * 16 < nb) {
    for (i = 0; i < na; ++i)
      if (bsearch_u32(b, nb, a[i]) >= 0) out[n++] = a[i];
    return n;
  }
  if (nb * 16 < na) {
    for (j = 0; j < nb; ++j)
      if (bsearch_u32(a, na, b[j]) >= 0) out[n++] = b[j];
    return n;
  }
  while (i < na && j < nb) {
    if (a[i] == b[j]) {
      out[n++] = a[i];
      ++i;
      ++j;
    } else if (a[i] < b[j])
      ++i;
    else
      ++j;
  }
  return n;
}
static int64_t intersect_cnt(const uint32_t* a, int64_t na, const uint32_t* b, int64_t nb) {
  int64_t i = 0, j = 0, n = 0;
  if (na * 16 < nb) {
    for (i = 0; i < na; ++i) n += bsearch_u32(b, nb, a[i]) >= 0;
    return n;
  }
  if (nb * 16 < na) {
    for (j = 0; j < nb; ++j) n += bsearch_u32(a, na, b[j]) >= 0;
    return n;
  }
  while (i < na && j < nb) {
    if (a[i] == b[j]) {
      ++n;
      ++i;
      ++j;
    } else if (a[i] < b[j])
      ++i;
    else
      ++j;
  }
  return n;
}
static int cmp_u32(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : x > y;
}

/* ------------------------------------------------------------------------ */
/* GMBE-order restatement                                                    */

typedef struct {
  int bound_height, bound_size; /* < 0: never split */
  int prune;
  int max_prefix; /* reference: 4 (LongTask holds 4 vertices) */
} gm_params;

typedef struct {
  const og* g;
  gm_params p;
  uint32_t *L, *Lp, *T1, *T2; /* dmaxV */
  int32_t* Llev;               /* dmaxV */
  int32_t* stk;                /* dmaxV + 2 */
  int64_t* lsz;                /* dmaxV + 2 */
  int32_t* cmap;               /* nV, zero between tasks */
  int32_t* vidx;               /* nV, zero between tasks (index + 1) */
  uint32_t *touched, *C, *R, *Rp; /* nV */
  int32_t *cnt, *lvl;          /* nV */
  uint32_t* Ltmp;              /* dmaxV, list-mode L' */
  uint32_t* Rtmp;              /* nV, list-mode R' */
  orc_result res;
  orc_list* list;
} gm_ws;

static int gm_ws_init(gm_ws* w, const og* g, gm_params p, int want_list) {
  memset(w, 0, sizeof(*w));
  w->g = g;
  w->p = p;
  size_t d = (size_t)g->dmaxV + 2, n = (size_t)g->nV + 2;
  w->L = malloc(d * 4);
  w->Lp = malloc(d * 4);
  w->T1 = malloc(d * 4);
  w->T2 = malloc(d * 4);
  w->Llev = malloc(d * 4);
  w->stk = malloc((d + 1) * 4);
  w->lsz = malloc((d + 1) * 8);
  w->Ltmp = malloc(d * 4);
  w->cmap = calloc(n, 4);
  w->vidx = calloc(n, 4);
  w->touched = malloc(n * 4);
  w->C = malloc(n * 4);
  w->R = malloc(n * 4);
  w->Rp = malloc(n * 4);
  w->cnt = malloc(n * 4);
  w->lvl = malloc(n * 4);
  w->Rtmp = malloc(n * 4);
  if (want_list) w->list = orc_list_new();
  return (w->L && w->Lp && w->T1 && w->T2 && w->Llev && w->stk && w->lsz && w->Ltmp && w->cmap &&
          w->vidx && w->touched && w->C && w->R && w->Rp && w->cnt && w->lvl && w->Rtmp)
             ? 0
             : -1;
}
static void gm_ws_free(gm_ws* w) {
  free(w->L); free(w->Lp); free(w->T1); free(w->T2); free(w->Llev); free(w->stk);
  free(w->lsz); free(w->Ltmp); free(w->cmap); free(w->vidx); free(w->touched);
  free(w->C); free(w->R); free(w->Rp); free(w->cnt); free(w->lvl); free(w->Rtmp);
  orc_list_free(w->list);
}

#define NV_ROW(g, v) ((g)->v_adj + (g)->v_off[v])
#define NV_DEG(g, v) ((g)->v_off[(v) + 1] - (g)->v_off[v])
#define NU_ROW(g, u) ((g)->u_adj + (g)->u_off[u])
#define NU_DEG(g, u) ((g)->u_off[(u) + 1] - (g)->u_off[u])

/* common neighbours of S[0..k) over V->U rows (get_common_neighbors,
 * src/IterFinderGpu.cu:680-711).  Writes the intersection of S[0..k-1) to Lp
 * (when k > 1) and of all k to out. */
static int64_t gm_common(gm_ws* w, const uint32_t* S, int k, uint32_t* out, int64_t* np) {
  const og* g = w->g;
  int64_t n = NV_DEG(g, S[0]);
  memcpy(w->T1, NV_ROW(g, S[0]), (size_t)n * 4);
  uint32_t *cur = w->T1, *nxt = w->T2;
  for (int i = 1; i < k; ++i) {
    if (i == k - 1) {
      memcpy(w->Lp, cur, (size_t)n * 4);
      *np = n;
    }
    n = intersect(cur, n, NV_ROW(g, S[i]), NV_DEG(g, S[i]), nxt);
    uint32_t* t = cur;
    cur = nxt;
    nxt = t;
  }
  memcpy(out, cur, (size_t)n * 4);
  return n;
}

/* count |N(w) ∩ L| for every w in the union of N(u), u in L; returns the
 * union (sorted) in w->touched with counts in w->cmap */
static int64_t gm_count(gm_ws* w, const uint32_t* L, int64_t nL) {
  const og* g = w->g;
  int64_t nt = 0;
  for (int64_t j = 0; j < nL; ++j) {
    const uint32_t* row = NU_ROW(g, L[j]);
    int64_t d = NU_DEG(g, L[j]);
    for (int64_t e = 0; e < d; ++e)
      if (w->cmap[row[e]]++ == 0) w->touched[nt++] = row[e];
  }
  qsort(w->touched, (size_t)nt, 4, cmp_u32);
  return nt;
}
static void gm_uncount(gm_ws* w, int64_t nt) {
  for (int64_t i = 0; i < nt; ++i) w->cmap[w->touched[i]] = 0;
}

/* finder_push (src/IterFinderGpu.h:206-264).  Returns maximality; *nnew gets
 * |L'|; on a maximal push the biclique is hashed / recorded here. */
static int gm_push(gm_ws* w, int cid, int level, int64_t nL, int64_t nC, uint64_t sumR0,
                   int64_t nR0, int64_t* nnew) {
  const og* g = w->g;
  uint32_t x = w->C[cid];
  const uint32_t* nx = NV_ROW(g, x);
  int64_t dx = NV_DEG(g, x);
  int64_t old = 0, nw = 0;
  for (int64_t j = 0; j < nL; ++j)
    if (w->Llev[j] == level - 1) {
      ++old;
      if (bsearch_u32(nx, dx, w->L[j]) >= 0) {
        w->Llev[j] = level;
        ++nw;
      }
    }
  int recompute = old > 2 * nw;
  int victim = recompute ? level : level - 1;
  if (recompute)
    for (int64_t i = 0; i < nC; ++i) w->cnt[i] = 0;
  for (int64_t j = 0; j < nL; ++j)
    if (w->Llev[j] == victim) {
      const uint32_t* row = NU_ROW(g, w->L[j]);
      int64_t d = NU_DEG(g, w->L[j]);
      for (int64_t e = 0; e < d; ++e) {
        int32_t p = w->vidx[row[e]];
        if (p) w->cnt[p - 1] += recompute ? 1 : -1;
      }
    }
  *nnew = nw;
  uint64_t sumR = sumR0;
  int64_t nR = nR0;
  if (w->list) memcpy(w->Rtmp, w->R, (size_t)nR0 * 4);
  for (int64_t i = 0; i < nC; ++i) {
    if (w->cnt[i] == nw) {
      if (i < cid && w->lvl[i] >= level) return 0;
      if (w->lvl[i] > level) w->lvl[i] = level;
      sumR += g->kV[w->C[i]];
      if (w->list) w->Rtmp[nR++] = w->C[i];
    }
  }
  uint64_t sumL = 0;
  int64_t k = 0;
  for (int64_t j = 0; j < nL; ++j)
    if (w->Llev[j] == level) {
      sumL += g->kU[w->L[j]];
      if (w->list) w->Ltmp[k++] = w->L[j];
    }
  add_biclique(&w->res, hash_internal(g, sumL, sumR));
  if (w->list) {
    qsort(w->Rtmp, (size_t)nR, 4, cmp_u32);
    record(g, w->list, w->Ltmp, (size_t)k, w->Rtmp, (size_t)nR);
  }
  return 1;
}

/* finder_pop (src/IterFinderGpu.h:266-303) */
static void gm_pop(gm_ws* w, int cid, int level, int64_t nL, int64_t nC) {
  const og* g = w->g;
  for (int64_t j = 0; j < nL; ++j) {
    if (w->Llev[j] == level - 1) {
      const uint32_t* row = NU_ROW(g, w->L[j]);
      int64_t d = NU_DEG(g, w->L[j]);
      for (int64_t e = 0; e < d; ++e) {
        int32_t p = w->vidx[row[e]];
        if (p) {
          w->cnt[p - 1]++;
          if (w->lvl[p - 1] >= level) w->lvl[p - 1] = LV_TOUCHED;
        }
      }
    } else if (w->Llev[j] >= level) {
      w->Llev[j] = level - 1;
    }
  }
  if (w->p.prune) {
    for (int64_t i = 0; i < nC; ++i)
      if (w->lvl[i] >= level)
        w->lvl[i] = (w->lvl[i] == LV_TOUCHED || i == cid) ? LV_ACTIVE : level - 1;
  } else {
    w->lvl[cid] = LV_ACTIVE;
  }
}

/* IterProcess (src/IterFinderGpu.cu:12-69) */
static void gm_iterate(gm_ws* w, int64_t nL, int64_t nC, int first_id, uint64_t sumR0, int64_t nR0,
                       int count_bytes) {
  const og* g = w->g;
  int level = 1;
  w->stk[0] = first_id;
  w->lsz[0] = nL;
  while (level != 0) {
    int cur = -1;
    for (int64_t i = w->stk[level - 1]; i < nC; ++i)
      if (w->lvl[i] > level - 1 && w->cnt[i] != 0) {
        cur = (int)i;
        break;
      }
    if (cur != -1) {
      int64_t nw = 0;
      int maximal = gm_push(w, cur, level, nL, nC, sumR0, nR0, &nw);
      w->res.nodes++;
      if (count_bytes) w->res.b_alg += 4u * (uint64_t)(NV_DEG(g, w->C[cur]) + w->lsz[level - 1] + nC);
      if (!maximal) w->res.non_maximal++;
      if (maximal && w->cnt[cur] > 1) {
        w->stk[level - 1] = cur;
        w->stk[level] = cur + 1;
        w->lsz[level] = nw;
        level++;
      } else {
        w->stk[level - 1] = cur + 1;
        gm_pop(w, cur, level, nL, nC);
      }
    } else {
      level--;
      if (level > 0) {
        cur = w->stk[level - 1];
        gm_pop(w, cur, level, nL, nC);
        w->stk[level - 1] = cur + 1;
      }
    }
  }
}

/* IterFinderWithMultipleVertex (src/IterFinderGpu.cu:772-887).  S holds k
 * ascending vertices; `top` marks a top-level vertex (for B_alg). */
static void gm_task(gm_ws* w, const uint32_t* S, int k) {
  const og* g = w->g;
  w->res.tasks++;
  int64_t np = 0;
  int64_t nL = gm_common(w, S, k, w->L, &np);
  int count_bytes = (k == 1) && w->p.bound_height < 0;
  if (count_bytes) {
    uint64_t b = (uint64_t)NV_DEG(g, S[0]);
    for (int64_t e = 0; e < NV_DEG(g, S[0]); ++e) b += (uint64_t)NU_DEG(g, NV_ROW(g, S[0])[e]);
    w->res.b_alg += 4u * b;
  }
  w->res.nodes++;
  if (nL == 0) { /* isolated vertex: no biclique (reference: UB, Q5) */
    w->res.non_maximal++;
    return;
  }
  uint32_t vk = S[k - 1];
  /* canonicity (src/IterFinderGpu.cu:800-816): R below v_k must equal the
   * closure of the prefix below v_k */
  int64_t nRp = 0;
  if (k > 1) {
    int64_t nt = gm_count(w, w->Lp, np);
    for (int64_t i = 0; i < nt && w->touched[i] < vk; ++i)
      if (w->cmap[w->touched[i]] == np) w->Rp[nRp++] = w->touched[i];
    gm_uncount(w, nt);
  }
  int64_t nt = gm_count(w, w->L, nL);
  int64_t nR = 0;
  for (int64_t i = 0; i < nt; ++i)
    if (w->cmap[w->touched[i]] == nL) w->R[nR++] = w->touched[i];
  int canonical = 1;
  if (k == 1) {
    canonical = nR > 0 && w->R[0] == S[0];
  } else {
    int64_t i = 0;
    for (; i < nR && w->R[i] < vk; ++i)
      if (i >= nRp || w->R[i] != w->Rp[i]) {
        canonical = 0;
        break;
      }
    if (canonical && i < nRp) canonical = 0;
  }
  if (!canonical) {
    w->res.non_maximal++;
    gm_uncount(w, nt);
    return;
  }
  uint64_t sumL = 0, sumR = 0;
  for (int64_t j = 0; j < nL; ++j) sumL += g->kU[w->L[j]];
  for (int64_t i = 0; i < nR; ++i) sumR += g->kV[w->R[i]];
  add_biclique(&w->res, hash_internal(g, sumL, sumR));
  if (w->list) record(g, w->list, w->L, (size_t)nL, w->R, (size_t)nR);
  if (nL <= 1) {
    gm_uncount(w, nt);
    return;
  }
  /* C = union \ R, with counts; next_id = #C below v_k (src/IterFinderGpu.cu:830-834) */
  int64_t nC = 0, next_id = 0;
  for (int64_t i = 0; i < nt; ++i) {
    uint32_t x = w->touched[i];
    int32_t c = w->cmap[x];
    if (c == nL) continue;
    if (x < vk) next_id++;
    w->C[nC] = x;
    w->cnt[nC] = c;
    w->lvl[nC] = LV_ACTIVE;
    nC++;
  }
  gm_uncount(w, nt);
  /* split rule (src/IterFinderGpu.cu:835-858); children as in LargeToTiny
   * (src/IterFinderGpu.cu:889-910) */
  int64_t rest = nC - next_id;
  if (w->p.bound_height >= 0 && k < w->p.max_prefix &&
      (rest > w->p.bound_height || nL > w->p.bound_height) &&
      rest * nL > (int64_t)w->p.bound_size) {
    w->res.splits++;
    uint32_t* kids = (uint32_t*)malloc((size_t)(rest ? rest : 1) * 4);
    memcpy(kids, w->C + next_id, (size_t)rest * 4);
    uint32_t S2[16];
    memcpy(S2, S, (size_t)k * 4);
    for (int64_t i = 0; i < rest; ++i) {
      S2[k] = kids[i];
      gm_task(w, S2, k + 1);
    }
    free(kids);
    return;
  }
  for (int64_t i = 0; i < nC; ++i) w->vidx[w->C[i]] = (int32_t)i + 1;
  for (int64_t j = 0; j < nL; ++j) w->Llev[j] = 0;
  gm_iterate(w, nL, nC, (int)next_id, sumR, nR, count_bytes);
  for (int64_t i = 0; i < nC; ++i) w->vidx[w->C[i]] = 0;
  /* the stack machine's R' for list mode starts from R: keep R intact */
}

/* ------------------------------------------------------------------------ */
/* MBEA restatement                                                          */

typedef struct {
  const og* g;
  int32_t* cmap;
  uint32_t* touched;
  uint32_t* arena;
  size_t arena_n, arena_cap;
  orc_result res;
  orc_list* list;
  double deadline; /* 0 = none; past it, subtrees are abandoned (counts partial) */
  uint64_t tick;
  int cut;
  const int32_t* range; /* PMBE: released-children range per V vertex, 2 ints */
} mb_ws;

static uint32_t* mb_alloc(mb_ws* w, size_t n) {
  if (w->arena_n + n > w->arena_cap) return NULL;
  uint32_t* p = w->arena + w->arena_n;
  w->arena_n += n;
  return p;
}

/* MbeaFinder::biclique_find (src/BicliqueFinder.cu:32-57) */
static int mb_find(mb_ws* w, const uint32_t* L, int64_t nL, const uint32_t* R, int64_t nR,
                   uint64_t sumR, const uint32_t* C, int64_t nC, uint32_t vc) {
  const og* g = w->g;
  if (w->deadline > 0 && (++w->tick & 255u) == 0 && wall() > w->deadline) w->cut = 1;
  if (w->cut) return 0;
  size_t mark = w->arena_n;
  uint32_t* Lp = mb_alloc(w, (size_t)nL + 1);
  if (!Lp) return -1;
  int64_t nLp = intersect(L, nL, NV_ROW(g, vc), NV_DEG(g, vc), Lp);
  w->res.nodes++;
  if (nLp == 0) {
    w->arena_n = mark;
    return 0;
  }
  uint32_t* Rp = mb_alloc(w, (size_t)(nR + nC) + 1);
  uint32_t* Cp = mb_alloc(w, (size_t)nC + 1);
  if (!Rp || !Cp) return -1;
  memcpy(Rp, R, (size_t)nR * 4);
  int64_t nRp = nR, nCp = 0;
  uint64_t sR = sumR;
  for (int64_t i = 0; i < nC; ++i) {
    uint32_t c = C[i];
    int64_t nc = intersect_cnt(Lp, nLp, NV_ROW(g, c), NV_DEG(g, c));
    if (nc == nLp) {
      if (c < vc) {
        w->res.non_maximal++;
        w->arena_n = mark;
        return 0;
      }
      Rp[nRp++] = c;
      sR += g->kV[c];
    } else if (nc != 0) {
      Cp[nCp++] = c;
    }
  }
  uint64_t sL = 0;
  for (int64_t j = 0; j < nLp; ++j) sL += g->kU[Lp[j]];
  add_biclique(&w->res, hash_internal(g, sL, sR));
  if (w->list) {
    qsort(Rp, (size_t)nRp, 4, cmp_u32);
    record(g, w->list, Lp, (size_t)nLp, Rp, (size_t)nRp);
  }
  for (int64_t i = 0; i < nCp; ++i)
    if (Cp[i] > vc)
      if (mb_find(w, Lp, nLp, Rp, nRp, sR, Cp, nCp, Cp[i]) < 0) return -1;
  w->arena_n = mark;
  return 0;
}

/* top level: 2-hop candidate set (MbeaAdvFinder::Execute,
 * baselines/MBE.zip:MBE/BicliqueFinder.cpp:42-88) feeding MbeaFinder's
 * recursion; equivalent to biclique_find(all U, {}, all V, v). */
static int mb_root(mb_ws* w, uint32_t v) {
  const og* g = w->g;
  const uint32_t* L = NV_ROW(g, v);
  int64_t nL = NV_DEG(g, v);
  w->res.tasks++;
  w->res.nodes++;
  if (nL == 0) return 0;
  int64_t nt = 0;
  for (int64_t j = 0; j < nL; ++j) {
    const uint32_t* row = NU_ROW(g, L[j]);
    int64_t d = NU_DEG(g, L[j]);
    for (int64_t e = 0; e < d; ++e)
      if (w->cmap[row[e]]++ == 0) w->touched[nt++] = row[e];
  }
  int canonical = 1;
  for (int64_t i = 0; i < nt; ++i)
    if (w->cmap[w->touched[i]] == nL && w->touched[i] < v) canonical = 0;
  if (!canonical) {
    w->res.non_maximal++;
    for (int64_t i = 0; i < nt; ++i) w->cmap[w->touched[i]] = 0;
    return 0;
  }
  qsort(w->touched, (size_t)nt, 4, cmp_u32);
  size_t mark = w->arena_n;
  uint32_t* R = mb_alloc(w, (size_t)nt + 1);
  uint32_t* C = mb_alloc(w, (size_t)nt + 1);
  if (!R || !C) return -1;
  int64_t nR = 0, nC = 0;
  uint64_t sR = 0, sL = 0;
  for (int64_t i = 0; i < nt; ++i) {
    uint32_t x = w->touched[i];
    if (w->cmap[x] == nL) {
      R[nR++] = x;
      sR += g->kV[x];
    } else {
      C[nC++] = x;
    }
    w->cmap[x] = 0;
  }
  for (int64_t j = 0; j < nL; ++j) sL += g->kU[L[j]];
  add_biclique(&w->res, hash_internal(g, sL, sR));
  if (w->list) record(g, w->list, L, (size_t)nL, R, (size_t)nR);
  for (int64_t i = 0; i < nC; ++i)
    if (C[i] > v)
      if (mb_find(w, L, nL, R, nR, sR, C, nC, C[i]) < 0) return -1;
  w->arena_n = mark;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* iMBEA restatement (the authors' improved CPU finder, -s 1 of MBE.zip)     */

typedef struct {
  uint32_t id;
  uint32_t pos; /* index in the parent's P */
  int64_t cnt;
} im_pair;

/* top level: overlap descending, then id descending (the comparator of
 * baselines/MBE.zip:MBE/BicliqueFinder.cpp:268-272, a total order) */
static int im_by_count_desc(const void* a, const void* b) {
  const im_pair* x = (const im_pair*)a;
  const im_pair* y = (const im_pair*)b;
  if (x->cnt != y->cnt) return x->cnt > y->cnt ? -1 : 1;
  return x->id > y->id ? -1 : (x->id < y->id ? 1 : 0);
}

/* inside the search: overlap descending only (BicliqueFinder.cpp:329-332);
 * std::sort leaves equal overlaps in an unspecified order, and keeps P's
 * order for up to 16 elements (libstdc++ sorts those by insertion): ties
 * keep P's order here */
static int im_by_count_desc_stable(const void* a, const void* b) {
  const im_pair* x = (const im_pair*)a;
  const im_pair* y = (const im_pair*)b;
  if (x->cnt != y->cnt) return x->cnt > y->cnt ? -1 : 1;
  return x->pos < y->pos ? -1 : (x->pos > y->pos ? 1 : 0);
}

/* n pairs from the u32 arena, 8-byte aligned */
static im_pair* im_pairs(mb_ws* w, size_t n) {
  uint32_t* p = mb_alloc(w, 4 * (n + 1) + 2);
  if (!p) return NULL;
  uintptr_t a = ((uintptr_t)p + 7u) & ~(uintptr_t)7u;
  return (im_pair*)a;
}

/* a \ b for ascending lists */
static int64_t except(const uint32_t* a, int64_t na, const uint32_t* b, int64_t nb, uint32_t* out) {
  int64_t i = 0, j = 0, n = 0;
  while (i < na) {
    while (j < nb && b[j] < a[i]) ++j;
    if (j < nb && b[j] == a[i]) {
      ++i;
      ++j;
    } else {
      out[n++] = a[i++];
    }
  }
  return n;
}

/* ImbeaAdvFinder::biclique_find (baselines/MBE.zip:MBE/BicliqueFinder.cpp:
 * 282-350).  P is processed from its back; a candidate whose neighbourhood
 * holds L' and nothing of L \ L' joins Q with x after x's subtree, and
 * leaves P as the reference's merge walk lets it (below).  P and Q are this
 * call's own arrays; Q has room for every entry of P. */
static int im_find(mb_ws* w, const uint32_t* L, int64_t nL, const uint32_t* R, int64_t nR,
                   uint64_t sR, uint32_t* P, int64_t nP, uint32_t* Q, int64_t nQ) {
  const og* g = w->g;
  const int64_t nQ0 = nQ; /* Q[nQ0..) holds this call's own additions */
  while (nP > 0) {
    if (w->deadline > 0 && (++w->tick & 255u) == 0 && wall() > w->deadline) w->cut = 1;
    if (w->cut) return 0;
    size_t mark = w->arena_n;
    uint32_t x = P[--nP];
    uint32_t* Lp = mb_alloc(w, (size_t)nL + 1);
    uint32_t* Lc = mb_alloc(w, (size_t)nL + 1);
    if (!Lp || !Lc) return -1;
    int64_t nLp = intersect(L, nL, NV_ROW(g, x), NV_DEG(g, x), Lp);
    int64_t nLc = except(L, nL, Lp, nLp, Lc);
    if (nLp == 0) {
      Q[nQ++] = x;
      w->arena_n = mark;
      continue;
    }
    w->res.nodes++;
    uint32_t* Qp = mb_alloc(w, (size_t)(nQ + nP) + 1);
    if (!Qp) return -1;
    int64_t nQp = 0;
    int maximal = 1;
    for (int64_t i = 0; i < nQ && maximal; ++i) {
      int64_t nc = intersect_cnt(Lp, nLp, NV_ROW(g, Q[i]), NV_DEG(g, Q[i]));
      if (nc == nLp) maximal = 0;
      else if (nc > 0) Qp[nQp++] = Q[i];
    }
    if (!maximal) {
      w->res.non_maximal++;
      w->arena_n = mark;
      continue;
    }
    uint32_t* Rp = mb_alloc(w, (size_t)(nR + nP) + 2);
    uint32_t* Cx = mb_alloc(w, (size_t)nP + 2);
    im_pair* pp = im_pairs(w, (size_t)nP);
    if (!Rp || !Cx || !pp) return -1;
    memcpy(Rp, R, (size_t)nR * 4);
    int64_t nRp = nR, nCx = 0, npp = 0;
    Rp[nRp++] = x;
    uint64_t sRp = sR + g->kV[x];
    Cx[nCx++] = x;
    for (int64_t i = 0; i < nP; ++i) {
      const uint32_t p = P[i];
      int64_t nc = intersect_cnt(Lp, nLp, NV_ROW(g, p), NV_DEG(g, p));
      if (nc == nLp) {
        Rp[nRp++] = p;
        sRp += g->kV[p];
        if (intersect_cnt(Lc, nLc, NV_ROW(g, p), NV_DEG(g, p)) == 0) Cx[nCx++] = p;
      } else if (nc > 0) {
        pp[npp].id = p;
        pp[npp].pos = (uint32_t)i;
        pp[npp].cnt = nc;
        ++npp;
      }
    }
    if (npp > 0) {
      qsort(pp, (size_t)npp, sizeof(im_pair), im_by_count_desc_stable);
      uint32_t* Pp = mb_alloc(w, (size_t)npp + 1);
      if (!Pp) return -1;
      for (int64_t i = 0; i < npp; ++i) Pp[i] = pp[i].id;
      if (im_find(w, Lp, nLp, Rp, nRp, sRp, Pp, npp, Qp, nQp) < 0) return -1;
    }
    uint64_t sL = 0;
    for (int64_t j = 0; j < nLp; ++j) sL += g->kU[Lp[j]];
    add_biclique(&w->res, hash_internal(g, sL, sRp));
    if (w->list) {
      qsort(Rp, (size_t)nRp, 4, cmp_u32);
      record(g, w->list, Lp, (size_t)nLp, Rp, (size_t)nRp);
    }
    /* P = seq_except(P, C) as the reference computes it
     * (BicliqueFinder.cpp:345, Utility.cpp:97-113): a merge walk that assumes
     * both lists ascend, applied to P in overlap order and C in P's order.
     * It drops only members of C and may keep some; a kept one is rejected
     * as non-maximal when it comes up (x, its twin on L, is in Q), so only
     * the node count follows the walk, and the result does not. */
    int64_t k = 0, c = 0;
    for (int64_t i = 0; i < nP;) {
      if (c == nCx || P[i] < Cx[c]) {
        P[k++] = P[i++];
      } else if (P[i] == Cx[c]) {
        ++i;
        ++c;
      } else {
        ++c;
      }
    }
    nP = k;
    /* Q += C.  A member kept in P can be in a later C again; Q holds it once
     * (a repeat changes no maximality test), so Q never outgrows P. */
    for (int64_t j = 0; j < nCx; ++j) {
      int seen = 0;
      for (int64_t q = nQ0; q < nQ && !seen; ++q) seen = Q[q] == Cx[j];
      if (!seen) Q[nQ++] = Cx[j];
    }
    w->arena_n = mark;
  }
  return 0;
}

/* ImbeaAdvFinder::Execute, one top-level vertex (baselines/MBE.zip:MBE/
 * BicliqueFinder.cpp:227-280): 2-hop counts, canonicity against lower ids,
 * the root biclique, then the search over the higher 2-hop neighbours
 * ordered by decreasing overlap. */
static int im_root(mb_ws* w, uint32_t v) {
  const og* g = w->g;
  const uint32_t* L = NV_ROW(g, v);
  int64_t nL = NV_DEG(g, v);
  w->res.tasks++;
  if (nL == 0) return 0;
  w->res.nodes++;
  int64_t nt = 0;
  for (int64_t j = 0; j < nL; ++j) {
    const uint32_t* row = NU_ROW(g, L[j]);
    int64_t d = NU_DEG(g, L[j]);
    for (int64_t e = 0; e < d; ++e)
      if (w->cmap[row[e]]++ == 0) w->touched[nt++] = row[e];
  }
  int canonical = 1;
  for (int64_t i = 0; i < nt; ++i)
    if (w->cmap[w->touched[i]] == nL && w->touched[i] < v) canonical = 0;
  if (!canonical) {
    w->res.non_maximal++;
    for (int64_t i = 0; i < nt; ++i) w->cmap[w->touched[i]] = 0;
    return 0;
  }
  size_t mark = w->arena_n;
  uint32_t* R = mb_alloc(w, (size_t)nt + 1);
  uint32_t* Q = mb_alloc(w, (size_t)nt + 1);
  uint32_t* P = mb_alloc(w, (size_t)nt + 1);
  im_pair* pp = im_pairs(w, (size_t)nt);
  if (!R || !Q || !P || !pp) return -1;
  int64_t nR = 0, nQ = 0, npp = 0;
  uint64_t sR = 0, sL = 0;
  for (int64_t i = 0; i < nt; ++i) {
    uint32_t y = w->touched[i];
    int64_t c = w->cmap[y];
    w->cmap[y] = 0;
    if (c == nL) {
      R[nR++] = y;
      sR += g->kV[y];
    } else if (y < v) {
      Q[nQ++] = y;
    } else {
      pp[npp].id = y;
      pp[npp].pos = 0;
      pp[npp].cnt = c;
      ++npp;
    }
  }
  qsort(R, (size_t)nR, 4, cmp_u32);
  for (int64_t j = 0; j < nL; ++j) sL += g->kU[L[j]];
  add_biclique(&w->res, hash_internal(g, sL, sR));
  if (w->list) record(g, w->list, L, (size_t)nL, R, (size_t)nR);
  qsort(pp, (size_t)npp, sizeof(im_pair), im_by_count_desc);
  for (int64_t i = 0; i < npp; ++i) P[i] = pp[i].id;
  int rc = im_find(w, L, nL, R, nR, sR, P, npp, Q, nQ);
  w->arena_n = mark;
  return rc;
}

/* ------------------------------------------------------------------------ */
/* PMBE restatement (the authors' CPU finder with the containment order,     */
/* -s 2 of MBE.zip)                                                          */

typedef struct {
  uint32_t deg, id;
} pm_key;

/* released vertices, larger neighbourhood first, then ascending id.  The
 * reference sorts them with std::sort by degree only, which leaves the order
 * of equal degrees unspecified; any order is a valid topological order and
 * yields the same bicliques. */
static int pm_by_deg_desc(const void* a, const void* b) {
  const pm_key* x = (const pm_key*)a;
  const pm_key* y = (const pm_key*)b;
  if (x->deg != y->deg) return x->deg > y->deg ? -1 : 1;
  return x->id < y->id ? -1 : (x->id > y->id ? 1 : 0);
}

/* PmbeBiGraph's vertex order (baselines/MBE.zip:MBE/PMBE.cpp:3-77).  y is a
 * parent of x when N(x) ⊆ N(y); of two equal neighbourhoods the lower id is
 * the parent (an empty one is the child of every non-empty one).  Kahn's
 * algorithm over that DAG hands out new ids from nV - 1 downwards; the
 * vertices one vertex releases get consecutive ids below those still queued,
 * and range[2 * new id] .. range[2 * new id + 1] holds them (first > second:
 * none).  rename[new id] = old id.  Containment is found by 2-hop counting
 * instead of the reference's all-pairs intersections; the DAG is the same. */
static int pm_order(uint32_t nU, uint32_t nV, const int64_t* v_off, const uint32_t* v_adj,
                    uint32_t* rename, int32_t* range) {
  int64_t E = v_off[nV];
  int64_t* u_off = calloc((size_t)nU + 1, sizeof(int64_t));
  uint32_t* u_adj = malloc((size_t)(E ? E : 1) * 4);
  int64_t* pos = malloc(((size_t)nU + 1) * sizeof(int64_t));
  int32_t* cnt = calloc((size_t)nV + 1, 4);
  uint32_t* touched = malloc(((size_t)nV + 1) * 4);
  int64_t* indeg = calloc((size_t)nV + 1, sizeof(int64_t));
  uint32_t* queue = malloc(((size_t)nV + 1) * 4);
  pm_key* ready = malloc(((size_t)nV + 1) * sizeof(pm_key));
  int rc = 0;
  if (!u_off || !u_adj || !pos || !cnt || !touched || !indeg || !queue || !ready) {
    rc = -1;
    goto out;
  }
  for (int64_t e = 0; e < E; ++e) u_off[v_adj[e] + 1]++;
  for (uint32_t u = 0; u < nU; ++u) u_off[u + 1] += u_off[u];
  memcpy(pos, u_off, ((size_t)nU + 1) * sizeof(int64_t));
  for (uint32_t v = 0; v < nV; ++v)
    for (int64_t e = v_off[v]; e < v_off[v + 1]; ++e) u_adj[pos[v_adj[e]]++] = v;
#define PM_DEG(v) (v_off[(v) + 1] - v_off[(v)])
  /* 2-hop counts of x: cnt[y] = |N(x) ∩ N(y)| over the y that share a neighbour */
#define PM_COUNT(x, nt)                                                  \
  do {                                                                   \
    nt = 0;                                                              \
    for (int64_t e = v_off[(x)]; e < v_off[(x) + 1]; ++e) {              \
      uint32_t u_ = v_adj[e];                                            \
      for (int64_t f = u_off[u_]; f < u_off[u_ + 1]; ++f)                \
        if (cnt[u_adj[f]]++ == 0) touched[nt++] = u_adj[f];              \
    }                                                                    \
  } while (0)
  int64_t n_full = 0, n_empty = 0;
  for (uint32_t x = 0; x < nV; ++x) {
    if (PM_DEG(x) == 0) {
      /* every non-empty vertex and every lower empty one is a parent */
      indeg[x] = -1 - n_empty++; /* fixed up below */
      continue;
    }
    ++n_full;
    int64_t nt;
    PM_COUNT(x, nt);
    for (int64_t i = 0; i < nt; ++i) {
      uint32_t y = touched[i];
      if (y != x && cnt[y] == PM_DEG(x) && (PM_DEG(y) > PM_DEG(x) || y < x)) indeg[x]++;
      cnt[y] = 0;
    }
  }
  for (uint32_t x = 0; x < nV; ++x)
    if (PM_DEG(x) == 0) indeg[x] = n_full + (-1 - indeg[x]);
  int64_t head = 0, tail = 0, nr = 0, done_full = 0;
  for (uint32_t x = 0; x < nV; ++x)
    if (indeg[x] == 0) ready[nr++] = (pm_key){(uint32_t)PM_DEG(x), x};
  qsort(ready, (size_t)nr, sizeof(pm_key), pm_by_deg_desc);
  while (nr > 0) queue[tail++] = ready[--nr].id;
  int64_t scan = nV;
  while (head < tail) {
    uint32_t r = queue[head++];
    rename[--scan] = r;
    int64_t start = scan - (tail - head), end = start - 1;
    nr = 0;
    if (PM_DEG(r) > 0) {
      int64_t nt;
      PM_COUNT(r, nt);
      for (int64_t i = 0; i < nt; ++i) {
        uint32_t y = touched[i];
        if (y != r && cnt[y] == PM_DEG(y) && (PM_DEG(y) < PM_DEG(r) || y > r))
          if (--indeg[y] == 0) ready[nr++] = (pm_key){(uint32_t)PM_DEG(y), y};
        cnt[y] = 0;
      }
      /* a non-empty vertex is a parent of every empty one; the last of them
       * releases the lowest empty id */
      if (++done_full == n_full && n_empty > 0)
        for (uint32_t x = 0; x < nV; ++x)
          if (PM_DEG(x) == 0) {
            ready[nr++] = (pm_key){0, x};
            break;
          }
    } else {
      /* an empty vertex releases the next empty id, its last parent */
      for (uint32_t x = r + 1; x < nV; ++x)
        if (PM_DEG(x) == 0) {
          ready[nr++] = (pm_key){0, x};
          break;
        }
    }
    start -= nr;
    qsort(ready, (size_t)nr, sizeof(pm_key), pm_by_deg_desc);
    while (nr > 0) queue[tail++] = ready[--nr].id;
    range[2 * scan] = (int32_t)start;
    range[2 * scan + 1] = (int32_t)end;
  }
  if (scan != 0) rc = -4; /* the order must place every vertex */
#undef PM_COUNT
#undef PM_DEG
out:
  free(u_off);
  free(u_adj);
  free(pos);
  free(cnt);
  free(touched);
  free(indeg);
  free(queue);
  free(ready);
  return rc;
}

/* common neighbours N(L) of a non-empty ascending L (U ids), ascending V ids */
static int64_t pm_common(const og* g, const uint32_t* L, int64_t nL, uint32_t* out) {
  int64_t n = NU_DEG(g, L[0]);
  memcpy(out, NU_ROW(g, L[0]), (size_t)n * 4);
  for (int64_t j = 1; j < nL && n > 0; ++j) {
    const uint32_t* row = NU_ROW(g, L[j]);
    int64_t d = NU_DEG(g, L[j]), k = 0, e = 0;
    for (int64_t i = 0; i < n; ++i) { /* in place: k <= i */
      while (e < d && row[e] < out[i]) ++e;
      if (e < d && row[e] == out[i]) out[k++] = out[i];
    }
    n = k;
  }
  return n;
}

/* PmbeAdvFinder::biclique_find (baselines/MBE.zip:MBE/PMBE.cpp:300-354).  C
 * ascends and is consumed from its back.  The range a consumed candidate
 * released joins this call's range queue; a candidate inside the range at the
 * queue's head is skipped, since its neighbourhood lies inside that of a
 * candidate already done.  A node is maximal when N(L') \ R lies within C. */
static int pm_find(mb_ws* w, const uint32_t* L, int64_t nL, const uint32_t* R, int64_t nR,
                   const uint32_t* C, int64_t nC) {
  const og* g = w->g;
  size_t base = w->arena_n;
  int32_t* ex = (int32_t*)mb_alloc(w, 2 * (size_t)nC + 2);
  if (!ex) return -1;
  int64_t qh = 0, qt = 0;
  while (nC > 0) {
    if (w->deadline > 0 && (++w->tick & 255u) == 0 && wall() > w->deadline) w->cut = 1;
    if (w->cut) break;
    uint32_t x = C[nC - 1];
    const int32_t* p = w->range + 2 * (size_t)x;
    if (p[0] <= p[1]) {
      ex[2 * qt] = p[0];
      ex[2 * qt + 1] = p[1];
      ++qt;
    }
    while (qh < qt && ex[2 * qh] > (int32_t)x) ++qh;
    if (qh < qt && ex[2 * qh + 1] >= (int32_t)x) {
      w->res.skipped++;
      --nC;
      continue;
    }
    size_t mark = w->arena_n;
    uint32_t* Lp = mb_alloc(w, (size_t)nL + 1);
    if (!Lp) return -1;
    int64_t nLp = intersect(L, nL, NV_ROW(g, x), NV_DEG(g, x), Lp);
    if (nLp == 0) {
      w->arena_n = mark;
      --nC;
      continue;
    }
    w->res.nodes++;
    uint32_t* Rp = mb_alloc(w, (size_t)NU_DEG(g, Lp[0]) + 1);
    if (!Rp) return -1;
    int64_t nRp = pm_common(g, Lp, nLp, Rp);
    uint32_t* Ra = mb_alloc(w, (size_t)nRp + 1);
    if (!Ra) return -1;
    int64_t nRa = except(Rp, nRp, R, nR, Ra);
    if (intersect_cnt(Ra, nRa, C, nC) != nRa) {
      w->res.non_maximal++;
    } else {
      uint32_t* Cp = mb_alloc(w, (size_t)nC + 1);
      if (!Cp) return -1;
      int64_t nCp = 0;
      for (int64_t i = 0; i < nC; ++i) {
        uint32_t c = C[i];
        if (c == x) continue;
        int64_t nc = intersect_cnt(Lp, nLp, NV_ROW(g, c), NV_DEG(g, c));
        if (nc > 0 && nc < nLp) Cp[nCp++] = c;
      }
      if (nCp > 0)
        if (pm_find(w, Lp, nLp, Rp, nRp, Cp, nCp) < 0) return -1;
      uint64_t sL = 0, sR = 0;
      for (int64_t j = 0; j < nLp; ++j) sL += g->kU[Lp[j]];
      for (int64_t j = 0; j < nRp; ++j) sR += g->kV[Rp[j]];
      add_biclique(&w->res, hash_internal(g, sL, sR));
      record(g, w->list, Lp, (size_t)nLp, Rp, (size_t)nRp);
    }
    w->arena_n = mark;
    --nC;
  }
  w->arena_n = base;
  return 0;
}

/* PmbeAdvFinder::Execute, one top-level vertex (baselines/MBE.zip:MBE/
 * PMBE.cpp:258-298): L = N(v), R = N(L); canonical when no lower id is in R;
 * then the search over the higher 2-hop neighbours not in R, ascending. */
static int pm_root(mb_ws* w, uint32_t v) {
  const og* g = w->g;
  const uint32_t* L = NV_ROW(g, v);
  int64_t nL = NV_DEG(g, v);
  w->res.tasks++;
  if (nL == 0) return 0;
  w->res.nodes++;
  int64_t nt = 0;
  for (int64_t j = 0; j < nL; ++j) {
    const uint32_t* row = NU_ROW(g, L[j]);
    int64_t d = NU_DEG(g, L[j]);
    for (int64_t e = 0; e < d; ++e)
      if (w->cmap[row[e]]++ == 0) w->touched[nt++] = row[e];
  }
  int canonical = 1;
  for (int64_t i = 0; i < nt; ++i)
    if (w->cmap[w->touched[i]] == nL && w->touched[i] < v) canonical = 0;
  if (!canonical) {
    w->res.non_maximal++;
    for (int64_t i = 0; i < nt; ++i) w->cmap[w->touched[i]] = 0;
    return 0;
  }
  qsort(w->touched, (size_t)nt, 4, cmp_u32);
  size_t mark = w->arena_n;
  uint32_t* R = mb_alloc(w, (size_t)nt + 1);
  uint32_t* C = mb_alloc(w, (size_t)nt + 1);
  if (!R || !C) return -1;
  int64_t nR = 0, nC = 0;
  uint64_t sR = 0, sL = 0;
  for (int64_t i = 0; i < nt; ++i) {
    uint32_t y = w->touched[i];
    if (w->cmap[y] == nL) {
      R[nR++] = y;
      sR += g->kV[y];
    } else if (y > v) {
      C[nC++] = y;
    }
    w->cmap[y] = 0;
  }
  for (int64_t j = 0; j < nL; ++j) sL += g->kU[L[j]];
  add_biclique(&w->res, hash_internal(g, sL, sR));
  record(g, w->list, L, (size_t)nL, R, (size_t)nR);
  int rc = pm_find(w, L, nL, R, nR, C, nC);
  w->arena_n = mark;
  return rc;
}

/* ------------------------------------------------------------------------ */
/* threading                                                                 */

typedef struct {
  int kind; /* 0 gmbe, 1 mbea, 2 imbea, 3 pmbe */
  const og* g;
  const int32_t* range; /* pmbe only */
  gm_params gp;
  int want_list;
  uint32_t stride, phase;
  double deadline; /* 0 = none */
  volatile int64_t next;  /* claim counter over the processing order */
  int64_t n_roots;
  volatile int failed;
} job;

typedef struct {
  job* j;
  orc_result res;
  orc_list* list;
  int rc;
} worker;

static int64_t claim(job* j) { return __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED); }

static void* run_worker(void* arg) {
  worker* wk = (worker*)arg;
  job* j = wk->j;
  const og* g = j->g;
  if (j->kind == 0) {
    gm_ws w;
    if (gm_ws_init(&w, g, j->gp, j->want_list)) {
      wk->rc = -1;
      j->failed = 1;
      gm_ws_free(&w);
      return NULL;
    }
    for (;;) {
      int64_t t = claim(j);
      if (t >= j->n_roots || j->failed) break;
      if (j->deadline > 0 && wall() > j->deadline) break;
      /* descending id (src/IterFinderGpu.cu:1008), restricted to the sample */
      uint32_t v = (uint32_t)(g->nV - 1 - ((uint64_t)t * j->stride + j->phase));
      gm_task(&w, &v, 1);
    }
    wk->res = w.res;
    if (w.list) {
      wk->list = w.list;
      w.list = NULL;
    }
    gm_ws_free(&w);
  } else {
    mb_ws w;
    memset(&w, 0, sizeof(w));
    w.g = g;
    w.cmap = calloc((size_t)g->nV + 1, 4);
    w.touched = malloc(((size_t)g->nV + 1) * 4);
    w.arena_cap = (size_t)1 << 26; /* 256 MB of u32 per thread */
    w.arena = malloc(w.arena_cap * 4);
    w.deadline = j->deadline;
    w.range = j->range;
    if (j->want_list) w.list = orc_list_new();
    if (!w.cmap || !w.touched || !w.arena) {
      wk->rc = -1;
      j->failed = 1;
    } else {
      for (;;) {
        int64_t t = claim(j);
        if (t >= j->n_roots || j->failed) break;
        if (j->deadline > 0 && wall() > j->deadline) break;
        uint32_t v = (uint32_t)(g->nV - 1 - ((uint64_t)t * j->stride + j->phase));
        int r = j->kind == 3 ? pm_root(&w, v) : (j->kind == 2 ? im_root(&w, v) : mb_root(&w, v));
        if (r < 0) {
          wk->rc = -2;
          j->failed = 1;
          break;
        }
      }
    }
    wk->res = w.res;
    wk->res.cut = (uint64_t)w.cut;
    wk->list = w.list;
    free(w.cmap);
    free(w.touched);
    free(w.arena);
  }
  return NULL;
}

static int run(int kind, uint32_t nU, uint32_t nV, const int64_t* v_off, const uint32_t* v_adj,
               const uint32_t* u_orig, const uint32_t* v_orig, int swapped, gm_params gp,
               int nthreads, uint32_t stride, uint32_t phase, double time_limit,
               const int32_t* range, orc_result* out, orc_list* list) {
  /* roots: processing ranks phase, phase + stride, ... (phase may exceed the
   * stride, which starts the sample at a later rank) */
  if (!out || !v_off || nthreads < 1 || stride < 1) return -1;
  memset(out, 0, sizeof(*out));
  og g;
  int rc = og_init(&g, nU, nV, v_off, v_adj, u_orig, v_orig, swapped);
  if (rc) {
    og_free(&g);
    return rc;
  }
  job j;
  memset(&j, 0, sizeof(j));
  j.kind = kind;
  j.g = &g;
  j.range = range;
  j.gp = gp;
  j.want_list = list != NULL;
  j.stride = stride;
  j.phase = phase;
  j.n_roots = nV > phase ? ((int64_t)nV - phase + stride - 1) / stride : 0;
  double t0 = wall();
  j.deadline = time_limit > 0 ? t0 + time_limit : 0;
  worker* wk = calloc((size_t)nthreads, sizeof(worker));
  pthread_t* th = calloc((size_t)nthreads, sizeof(pthread_t));
  pthread_attr_t attr;
  pthread_attr_init(&attr);
  pthread_attr_setstacksize(&attr, (size_t)256 << 20);
  for (int i = 0; i < nthreads; ++i) {
    wk[i].j = &j;
    pthread_create(&th[i], &attr, run_worker, &wk[i]);
  }
  for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
  pthread_attr_destroy(&attr);
  out->seconds = wall() - t0;
  int err = 0;
  int64_t claimed = j.next < j.n_roots ? j.next : j.n_roots;
  for (int i = 0; i < nthreads; ++i) {
    orc_result* r = &wk[i].res;
    out->maximal += r->maximal;
    out->non_maximal += r->non_maximal;
    out->hash_sum += r->hash_sum;
    out->hash_xor ^= r->hash_xor;
    out->nodes += r->nodes;
    out->b_alg += r->b_alg;
    out->tasks += r->tasks;
    out->splits += r->splits;
    out->skipped += r->skipped;
    out->cut += r->cut;
    if (wk[i].rc) err = wk[i].rc;
    if (wk[i].list) {
      if (list) list_append(list, wk[i].list);
      orc_list_free(wk[i].list);
    }
  }
  (void)claimed;
  free(wk);
  free(th);
  og_free(&g);
  return err;
}

/* Public entry points.  bound_height < 0 means the unsplit traversal (the
 * B_alg reference); max_prefix is the longest task prefix (reference: 4). */
int orc_gmbe(uint32_t nU, uint32_t nV, const int64_t* v_off, const uint32_t* v_adj,
             const uint32_t* u_orig, const uint32_t* v_orig, int swapped, int bound_height,
             int bound_size, int prune, int max_prefix, int nthreads, uint32_t stride,
             uint32_t phase, double time_limit, orc_result* out, orc_list* list) {
  gm_params p = {bound_height, bound_size, prune, max_prefix < 1 ? 4 : (max_prefix > 15 ? 15 : max_prefix)};
  return run(0, nU, nV, v_off, v_adj, u_orig, v_orig, swapped, p, nthreads, stride, phase,
             time_limit, NULL, out, list);
}

int orc_imbea(uint32_t nU, uint32_t nV, const int64_t* v_off, const uint32_t* v_adj,
              const uint32_t* u_orig, const uint32_t* v_orig, int swapped, int nthreads,
              uint32_t stride, uint32_t phase, double time_limit, orc_result* out,
              orc_list* list) {
  gm_params p = {-1, 0, 1, 4};
  return run(2, nU, nV, v_off, v_adj, u_orig, v_orig, swapped, p, nthreads, stride, phase,
             time_limit, NULL, out, list);
}

/* PMBE: V is renumbered by the containment order first; out->seconds times
 * the search only.  Processing rank t is the t-th vertex Kahn's algorithm
 * placed.  Hashes and lists carry original ids. */
int orc_pmbe(uint32_t nU, uint32_t nV, const int64_t* v_off, const uint32_t* v_adj,
             const uint32_t* u_orig, const uint32_t* v_orig, int swapped, int nthreads,
             uint32_t stride, uint32_t phase, double time_limit, orc_result* out,
             orc_list* list) {
  if (!out || !v_off) return -1;
  int64_t E = v_off[nV];
  uint32_t* rename = malloc(((size_t)nV + 1) * 4);
  int32_t* range = malloc(((size_t)nV + 1) * 8);
  int64_t* n_off = malloc(((size_t)nV + 1) * sizeof(int64_t));
  uint32_t* n_adj = malloc((size_t)(E ? E : 1) * 4);
  uint32_t* n_orig = malloc(((size_t)nV + 1) * 4);
  int rc = -1;
  if (rename && range && n_off && n_adj && n_orig) {
    for (uint32_t v = 0; v < nV; ++v)
      for (int64_t e = v_off[v]; e < v_off[v + 1]; ++e)
        if (v_adj[e] >= nU || (e > v_off[v] && v_adj[e] <= v_adj[e - 1])) {
          rc = -2;
          goto out;
        }
    rc = pm_order(nU, nV, v_off, v_adj, rename, range);
    if (rc == 0) {
      n_off[0] = 0;
      for (uint32_t i = 0; i < nV; ++i) {
        uint32_t o = rename[i];
        int64_t d = v_off[o + 1] - v_off[o];
        memcpy(n_adj + n_off[i], v_adj + v_off[o], (size_t)d * 4);
        n_off[i + 1] = n_off[i] + d;
        n_orig[i] = v_orig ? v_orig[o] : o;
      }
      gm_params p = {-1, 0, 1, 4};
      rc = run(3, nU, nV, n_off, n_adj, u_orig, n_orig, swapped, p, nthreads, stride, phase,
               time_limit, range, out, list);
    }
  }
out:
  free(rename);
  free(range);
  free(n_off);
  free(n_adj);
  free(n_orig);
  return rc;
}

int orc_mbea(uint32_t nU, uint32_t nV, const int64_t* v_off, const uint32_t* v_adj,
             const uint32_t* u_orig, const uint32_t* v_orig, int swapped, int nthreads,
             uint32_t stride, uint32_t phase, double time_limit, orc_result* out,
             orc_list* list) {
  gm_params p = {-1, 0, 1, 4};
  return run(1, nU, nV, v_off, v_adj, u_orig, v_orig, swapped, p, nthreads, stride, phase,
             time_limit, NULL, out, list);
}

/* Definition check: every subset R of V. */
int orc_brute(uint32_t nU, uint32_t nV, const int64_t* v_off, const uint32_t* v_adj,
              const uint32_t* u_orig, const uint32_t* v_orig, int swapped, orc_result* out,
              orc_list* list) {
  if (!out || nU > 64 || nV > 24) return -1;
  memset(out, 0, sizeof(*out));
  og g;
  int rc = og_init(&g, nU, nV, v_off, v_adj, u_orig, v_orig, swapped);
  if (rc) {
    og_free(&g);
    return rc;
  }
  uint64_t mask[24];
  for (uint32_t v = 0; v < nV; ++v) {
    mask[v] = 0;
    for (int64_t e = v_off[v]; e < v_off[v + 1]; ++e) mask[v] |= 1ull << v_adj[e];
  }
  uint32_t Lb[64], Rb[24];
  for (uint64_t s = 1; s < (1ull << nV); ++s) {
    uint64_t L = ~0ull;
    for (uint32_t v = 0; v < nV; ++v)
      if (s >> v & 1) L &= mask[v];
    if (nU < 64) L &= (1ull << nU) - 1;
    if (!L) continue;
    int maximal = 1;
    for (uint32_t v = 0; v < nV && maximal; ++v)
      if (!(s >> v & 1) && (mask[v] & L) == L) maximal = 0;
    if (!maximal) continue;
    uint64_t sL = 0, sR = 0;
    size_t nl = 0, nr = 0;
    for (uint32_t u = 0; u < nU; ++u)
      if (L >> u & 1) {
        sL += g.kU[u];
        Lb[nl++] = u;
      }
    for (uint32_t v = 0; v < nV; ++v)
      if (s >> v & 1) {
        sR += g.kV[v];
        Rb[nr++] = v;
      }
    add_biclique(out, hash_internal(&g, sL, sR));
    if (list) record(&g, list, Lb, nl, Rb, nr);
  }
  og_free(&g);
  return 0;
}
