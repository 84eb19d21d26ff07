// mbe_kernel.hpp — the wave64 maximal-biclique enumerator (device code).
//
// One wavefront runs one task at a time.  A task is a vertex prefix
// S = (v1 < ... < vk) of the enumeration side V (the reference's LongTask,
// src/GpuUtil.cuh:394-464): k = 1 for a top-level vertex, k > 1 for a child
// of a split task.  Per task (IterFinderWithMultipleVertex,
// src/IterFinderGpu.cu:772-887):
//   L  = common neighbours of S                         (wave intersections)
//   cnt(w) = |N(w) ∩ L| for w in the 2-hop set, sorted   (one edge pass with
//            L2 atomics into a per-wave count map + two-level bitmap)
//   R  = {w : cnt(w) = |L|}, canonical iff R below vk equals the closure of
//        the prefix below vk (src/IterFinderGpu.cu:800-816)
//   C  = 2-hop set \ R; next_id = |{c in C : c < vk}|
//   split into children S+[c] (c in C, c > vk) by the bound_height /
//        bound_size rule (src/IterFinderGpu.cu:835-858), or
//   enumerate the subtree with the stack machine below.
//
// Stack machine (IterProcess + finder_push/finder_pop,
// src/IterFinderGpu.cu:12-69, src/IterFinderGpu.h:206-303), redesigned: the
// reference keeps one candidate array with a level:16|count:16 word per
// candidate and rescans all of C at every node.  Here a node owns a compacted
// list of the candidates that still touch its L; each entry carries the
// candidate's neighbourhood inside L as a bitset over the task's L indices
// (W 64-bit words), so the local neighbourhood size is a popcount and
//   push x:  L' = a_x;  for every entry e of the parent list, m = a_e & a_x
//            m == a_x  -> e before x: x is not canonical (finder_push :252-261)
//                         e after x:  e joins R' (the level mark of :258-259)
//            0 < m < a_x -> e enters the child's list (branchable iff e is
//                         after x and still branchable)
//            a_e ⊆ a_x, e after x -> e is pruned at this node: PRUNE_EN's
//                         "untouched by the removed L vertices" rule (:291-299)
// is a single coalesced pass over the parent's list.  The pushes, the
// maximality verdicts and the pruning decisions are those of the reference.
// Lists live in a per-wave LDS arena when they fit, else in a per-wave HBM
// arena.
#pragma once

#include "wave.hpp"

namespace mbe {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kBranch = 0x80000000u;  // entry may still be pushed
constexpr int kMaxPrefix = 8;              // task capacity (vertices)
constexpr int kMaxW = 16;                  // widest bitset (|L| <= 1024)
constexpr int kMainW = 4;                  // widest bitset of the main kernel

struct Task {
  uint32_t v[kMaxPrefix];
};

struct DevGraph {
  const int64_t* v_off;
  const uint32_t* v_adj;
  const int64_t* u_off;
  const uint32_t* u_adj;
  const uint64_t* kU;
  const uint64_t* kV;
  uint32_t nU, nV, dmaxV, swapped;
};

// Counter slots of the result block (u64 each).
enum Acc : int {
  A_MAXIMAL = 0, A_NONMAX, A_HSUM, A_HXOR, A_NODES, A_TASKS, A_SPLITS, A_OVERFLOW,
  A_LIST_OVF, A_BALG, A_COUNT
};

struct Sched {
  unsigned long long* root_next;  // claim counter over top-level vertices
  unsigned long long* work_left;  // outstanding tasks (roots + children)
  uint32_t* q;                    // ring of kEntryWords-word entries
  unsigned long long* qseq;       // per-slot sequence numbers
  unsigned long long* qhead;      // next slot to pop
  unsigned long long* qtail;      // next slot to reserve for a push
  unsigned long long qcap;        // power of two
  unsigned long long n_roots;     // top-level vertices of this run (sample)
  uint32_t stride, phase;         // root r -> vertex nV-1-(r*stride+phase)
  const Task* root_tasks;         // if set, root r is root_tasks[r] instead
  Task* deferred;                 // tasks too wide for this kernel's bitsets
  unsigned long long* n_deferred;
  unsigned long long deferred_cap;
};

struct Scratch {
  char* base;        // per-wave slabs
  uint64_t slab;     // bytes per wave
  // offsets inside a slab
  uint64_t o_cmap, o_bm0, o_bm1, o_L, o_Lp, o_T, o_kL, o_U, o_Ucnt, o_C, o_kC, o_R, o_fr,
      o_arena;
  uint64_t arena_bytes;
  uint32_t nW0, nW1;
};

struct Params {
  DevGraph g;
  Sched s;
  Scratch w;
  unsigned long long* acc;   // A_COUNT counters
  uint32_t* list;            // list-mode output (internal ids) or null
  unsigned long long* list_cursor;
  unsigned long long list_cap;
  int bound_height, bound_size, prune, max_prefix, split;
};

struct Counters {
  uint64_t maximal = 0, nonmax = 0, hsum = 0, hxor = 0, nodes = 0, tasks = 0, splits = 0,
           overflow = 0;
  uint64_t balg = 0;  // algorithmic bytes: 4 B per id of every list a node must read
};

struct Frame {
  uint64_t sumR;
  uint64_t rec;      // generic address of the first record
  uint32_t len, cur, nbits, tier;  // tier 0 = LDS, 1 = HBM
  uint32_t top;      // arena offset of this allocation (freed on pop)
  uint32_t pad;
};

// ---------------------------------------------------------------------------
// helpers

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int64_t bsearch_u32(const uint32_t* a, int64_t n, uint32_t key) {
  int64_t lo = 0, hi = n - 1;
  while (lo <= hi) {
    int64_t mid = (lo + hi) >> 1;
    uint32_t x = a[mid];
    if (x == key) return mid;
    if (x < key) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

// out = A ∩ B (ascending); the smaller list is lane-strided and each lane
// binary-searches the larger one (seq_intersect_warp, src/GpuUtil.cuh:135-162).
__device__ uint32_t wave_intersect(const uint32_t* A, uint32_t nA, const uint32_t* B,
                                   uint32_t nB, uint32_t* out) {
  if (nA > nB) {
    const uint32_t* t = A; A = B; B = t;
    uint32_t tn = nA; nA = nB; nB = tn;
  }
  uint32_t cnt = 0, lane = lane_id();
  for (uint32_t base = 0; base < nA; base += kWave) {
    uint32_t i = base + lane;
    bool hit = false;
    uint32_t x = 0;
    if (i < nA) {
      x = A[i];
      hit = bsearch_u32(B, nB, x) >= 0;
    }
    uint64_t m = ballot(hit);
    if (hit) out[cnt + lanes_below(m)] = x;
    cnt += __popcll(m);
  }
  wave_sync();
  return cnt;
}

__device__ uint32_t wave_intersect_count(const uint32_t* A, uint32_t nA, const uint32_t* B,
                                         uint32_t nB) {
  if (nA > nB) {
    const uint32_t* t = A; A = B; B = t;
    uint32_t tn = nA; nA = nB; nB = tn;
  }
  uint32_t cnt = 0, lane = lane_id();
  for (uint32_t base = 0; base < nA; base += kWave) {
    uint32_t i = base + lane;
    bool hit = i < nA && bsearch_u32(B, nB, A[i]) >= 0;
    cnt += __popcll(ballot(hit));
  }
  return cnt;
}

// f(j, w) for every edge (L[j], w), w in N(L[j]) over the U->V rows.  The
// edges of 64 L vertices at a time are flattened across lanes; each lane
// finds its vertex by binary search over the 64 prefix offsets kept in LDS.
template <class F>
__device__ uint64_t for_each_edge(const uint32_t* L, uint32_t nL, const int64_t* off,
                                  const uint32_t* adj, uint32_t* s_off, int64_t* s_start, F&& f) {
  uint32_t lane = lane_id();
  uint64_t edges = 0;
  for (uint32_t base = 0; base < nL; base += kWave) {
    uint32_t j = base + lane;
    uint32_t deg = 0;
    int64_t st = 0;
    if (j < nL) {
      uint32_t u = L[j];
      st = off[u];
      deg = (uint32_t)(off[u + 1] - st);
    }
    uint32_t incl = wave_incl_scan32(deg);
    uint32_t total = read_lane(incl, kWave - 1);
    uint32_t nb = nL - base < (uint32_t)kWave ? nL - base : (uint32_t)kWave;
    edges += total;
    s_off[lane] = incl - deg;
    s_start[lane] = st;
    wave_sync();
    for (uint32_t t = lane; t < total; t += kWave) {
      uint32_t lo = 0, hi = nb - 1;
      while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (s_off[mid] <= t) lo = mid; else hi = mid - 1;
      }
      int64_t e = s_start[lo] + (int64_t)(t - s_off[lo]);
      f(base + lo, adj[e]);
    }
    wave_sync();
  }
  return edges;
}

// ---------------------------------------------------------------------------
// task queue.  Replaces WorkList<LargeTask> (src/WorkList.cuh:6-250) and the
// divider warp's LargeToTiny expansion (src/IterFinderGpu.cu:889-910,954-988):
// a split pushes its children directly, kChunk of them per 128-byte entry
// (the parent prefix + the children's vertex ids), so the number of global
// queue operations is about tasks / kChunk.  The ring is a bounded MPMC queue
// with per-slot sequence numbers; head and tail are reserved with one CAS,
// entries are written with device-scope stores and published by the slot's
// sequence number after a drain (no L2 writeback / invalidate is needed).

constexpr int kChunk = 24;      // children per entry
constexpr int kEntryWords = 32; // hdr, prefix[7], children[24]

__device__ __forceinline__ void backoff(uint32_t& n) {
  for (uint32_t i = 0; i < n; ++i) __builtin_amdgcn_s_sleep(32);
  n = n < 64 ? n * 2 : 64;
}

// Pop one entry into `word` (lane l < 32 holds word l).
__device__ bool queue_pop(const Sched& s, uint32_t& word) {
  uint32_t lane = lane_id();
  unsigned long long got = ~0ull;
  if (lane == 0) {
    unsigned long long h = ld_dev(s.qhead);
    for (;;) {
      unsigned long long tl = ld_dev(s.qtail);
      if (h >= tl) break;
      if (__hip_atomic_compare_exchange_strong(s.qhead, &h, h + 1, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        got = h;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    if (got != ~0ull) {
      unsigned long long slot = got & (s.qcap - 1);
      while (ld_dev(s.qseq + slot) != got + 1) __builtin_amdgcn_s_sleep(1);
    }
  }
  got = uni64(got);
  if (got == ~0ull) return false;
  wave_drain();
  unsigned long long slot = got & (s.qcap - 1);
  word = lane < (uint32_t)kEntryWords ? ld_dev(s.q + slot * kEntryWords + lane) : 0u;
  wave_drain();
  if (lane == 0) st_dev(s.qseq + slot, got + s.qcap);
  return true;
}

// Push the children prefix+[cand[i]], i < n, as ceil(n / kChunk) entries.
// All or nothing: false (nothing pushed) when the ring lacks room.
__device__ bool queue_push_children(const Sched& s, const Task& prefix, int k,
                                    const uint32_t* cand, uint32_t n, uint32_t& n_entries) {
  uint32_t lane = lane_id();
  const uint32_t ne = (n + kChunk - 1) / kChunk;
  n_entries = ne;
  unsigned long long start = ~0ull;
  if (lane == 0) {
    unsigned long long tl = ld_dev(s.qtail);
    for (;;) {
      unsigned long long h = ld_dev(s.qhead);
      if (tl + ne > h + s.qcap) break;
      if (__hip_atomic_compare_exchange_strong(s.qtail, &tl, tl + ne, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        start = tl;
        break;
      }
    }
  }
  start = uni64(start);
  if (start == ~0ull) return false;
  // two entries per pass: lanes 0-31 and 32-63
  for (uint32_t e0 = 0; e0 < ne; e0 += 2) {
    uint32_t e = e0 + (lane >> 5), w = lane & 31;
    if (e < ne) {
      unsigned long long pos = start + e, slot = pos & (s.qcap - 1);
      if (w == 0)
        while (ld_dev(s.qseq + slot) != pos) __builtin_amdgcn_s_sleep(1);
    }
    wave_drain();
    if (e < ne) {
      unsigned long long slot = (start + e) & (s.qcap - 1);
      uint32_t cnt = n - e * kChunk < (uint32_t)kChunk ? n - e * kChunk : (uint32_t)kChunk;
      uint32_t val;
      if (w == 0) val = (uint32_t)k | (cnt << 8);
      else if (w <= 7) val = w - 1 < (uint32_t)k ? prefix.v[w - 1] : kNone;
      else val = (w - 8) < cnt ? cand[e * kChunk + (w - 8)] : kNone;
      st_dev(s.q + slot * kEntryWords + w, val);
    }
    wave_drain();  // the words are in place before the slot is published
    if (e < ne && w == 0) st_dev(s.qseq + ((start + e) & (s.qcap - 1)), start + e + 1);
  }
  wave_drain();
  return true;
}

// Next unit of work: a queue entry (split children) first, then a top-level
// vertex in descending id (src/IterFinderGpu.cu:992-1018).  Returns 0 once
// every unit has finished (the gc/lc/tc protocol of
// src/IterFinderGpu.cu:978-986,1020-1034 collapsed into one outstanding-work
// counter), 1 for an entry (in `word`), 2 for a root task (in `t`, `k`).
__device__ int next_work(const Params& P, uint32_t& word, Task& t, int& k) {
  const Sched& s = P.s;
  uint32_t lane = lane_id();
  uint32_t nap = 1;
  for (;;) {
    if (queue_pop(s, word)) return 1;
    unsigned long long r = ~0ull;
    if (lane == 0 && ld_dev(s.root_next) < s.n_roots) {
      r = __hip_atomic_fetch_add(s.root_next, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    r = uni64(r);
    if (r != ~0ull && r < s.n_roots) {
      if (s.root_tasks) {
        uint32_t w = lane < (uint32_t)kMaxPrefix ? s.root_tasks[r].v[lane] : kNone;
        k = 0;
        for (int i = 0; i < kMaxPrefix; ++i) {
          t.v[i] = read_lane(w, i);
          k += t.v[i] != kNone;
        }
        return 2;
      }
      uint64_t rank = r * s.stride + s.phase;
      for (int i = 0; i < kMaxPrefix; ++i) t.v[i] = kNone;
      t.v[0] = P.g.nV - 1 - (uint32_t)rank;
      k = 1;
      return 2;
    }
    unsigned long long left = uni64(lane == 0 ? ld_dev(s.work_left) : 0ull);
    if (left == 0) return 0;
    backoff(nap);
  }
}

__device__ __forceinline__ void work_done(const Params& P) {
  if (lane_id() == 0)
    __hip_atomic_fetch_add(P.s.work_left, (unsigned long long)-1ll, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// per-wave state

struct WaveMem {
  int32_t* cmap;   // |V| counts / candidate index map, zero between tasks
  uint64_t* bm0;   // bit per V vertex touched
  uint64_t* bm1;   // bit per nonzero bm0 word
  uint32_t* L;     // task L (k > 1)
  uint32_t* Lp;    // prefix L
  uint32_t* T;     // temporary
  uint64_t* kL;    // hash keys of L
  uint32_t* U;     // 2-hop set, ascending
  uint32_t* Ucnt;  // |N(w) ∩ L|
  uint32_t* C;     // candidates
  uint64_t* kC;    // hash keys of C
  uint32_t* R;     // closure of L
  Frame* fr;       // DFS frames
  char* arena;     // HBM list arena
  uint64_t arena_bytes;
  char* lds;       // LDS list arena
  uint32_t lds_bytes;
  uint32_t* s_off; // 64 + 1 u32 in LDS
  int64_t* s_start;// 64 i64 in LDS
};

__device__ __forceinline__ WaveMem wave_mem(const Params& P, uint64_t gw, char* lds,
                                            uint32_t lds_bytes, uint32_t* s_off,
                                            int64_t* s_start) {
  const Scratch& S = P.w;
  char* b = S.base + gw * S.slab;
  WaveMem m;
  m.cmap = (int32_t*)(b + S.o_cmap);
  m.bm0 = (uint64_t*)(b + S.o_bm0);
  m.bm1 = (uint64_t*)(b + S.o_bm1);
  m.L = (uint32_t*)(b + S.o_L);
  m.Lp = (uint32_t*)(b + S.o_Lp);
  m.T = (uint32_t*)(b + S.o_T);
  m.kL = (uint64_t*)(b + S.o_kL);
  m.U = (uint32_t*)(b + S.o_U);
  m.Ucnt = (uint32_t*)(b + S.o_Ucnt);
  m.C = (uint32_t*)(b + S.o_C);
  m.kC = (uint64_t*)(b + S.o_kC);
  m.R = (uint32_t*)(b + S.o_R);
  m.fr = (Frame*)(b + S.o_fr);
  m.arena = b + S.o_arena;
  m.arena_bytes = S.arena_bytes;
  m.lds = lds;
  m.lds_bytes = lds_bytes;
  m.s_off = s_off;
  m.s_start = s_start;
  return m;
}

__device__ __forceinline__ uint64_t biclique_hash(uint64_t sumL, uint64_t sumR, uint32_t swapped) {
  uint64_t sa = swapped ? sumR : sumL, sb = swapped ? sumL : sumR;
  return mix64(sa ^ rotl64(sb, 31));
}
__device__ __forceinline__ void add_biclique(Counters& c, uint64_t h) {
  c.maximal++;
  c.hsum += h;
  c.hxor ^= h;
}

// List mode: reserve and write a record [nL, nR, L..., R...] (internal ids).
__device__ uint64_t list_reserve(const Params& P, uint64_t words) {
  unsigned long long at = ~0ull;
  if (lane_id() == 0) {
    at = __hip_atomic_fetch_add(P.list_cursor, (unsigned long long)words, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
    if (at + words > P.list_cap) {
      __hip_atomic_fetch_add(P.acc + A_LIST_OVF, 1ull, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      at = ~0ull;
    }
  }
  return uni64(at);
}

// ---------------------------------------------------------------------------
// the DFS over one task (W = bitset words per entry)

template <int W>
struct Rec {
  uint32_t idx;  // candidate index (into C) | kBranch
  uint32_t pad;
  uint64_t a[W];
};

// Allocate `bytes` for a child list: LDS first, then HBM.  Returns the
// generic address or 0 when neither arena has room.
__device__ __forceinline__ uint64_t arena_take(const WaveMem& m, uint32_t& lds_top,
                                               uint64_t& hbm_top, uint64_t bytes,
                                               uint32_t& tier, uint64_t& top_before) {
  if ((uint64_t)lds_top + bytes <= m.lds_bytes) {
    tier = 0;
    top_before = lds_top;
    return (uint64_t)(m.lds + lds_top);
  }
  if (hbm_top + bytes <= m.arena_bytes) {
    tier = 1;
    top_before = hbm_top;
    return (uint64_t)(m.arena + hbm_top);
  }
  return 0;
}

template <int W>
__device__ void dfs(const Params& P, const WaveMem& m, uint32_t nL, uint32_t nC,
                    uint32_t first, uint64_t sumR0, uint64_t root_rec, uint32_t root_tier,
                    uint64_t root_top, uint32_t& lds_top, uint64_t& hbm_top, Counters& ctr,
                    uint32_t nR0) {
  typedef Rec<W> R_t;
  const uint32_t lane = lane_id();
  const bool prune = P.prune != 0;
  Frame F;
  F.sumR = sumR0;
  F.rec = root_rec;
  F.len = nC;
  F.cur = first;
  F.nbits = nL;
  F.tier = root_tier;
  F.top = (uint32_t)root_top;
  uint32_t depth = 0;
  for (;;) {
    if (F.cur == kNone) {
      // frame exhausted: free its list, return to the parent
      if (F.tier == 0) lds_top = F.top; else hbm_top = F.top;
      if (depth == 0) break;
      --depth;
      wave_mem_wait();
      F = m.fr[depth];
      F.sumR = uni64(F.sumR);
      F.rec = uni64(F.rec);
      F.len = uni(F.len);
      F.cur = uni(F.cur);
      F.nbits = uni(F.nbits);
      F.tier = uni(F.tier);
      F.top = uni(F.top);
      continue;
    }
    R_t* recs = (R_t*)F.rec;
    const uint32_t xpos = F.cur;
    {
      // B(n) = 4 (deg(x) + |L_parent| + |C|)
      uint32_t xc = P.g.nV ? m.C[uni(recs[xpos].idx) & ~kBranch] : 0;
      uint64_t dx = (uint64_t)(P.g.v_off[xc + 1] - P.g.v_off[xc]);
      ctr.balg += 4ull * (dx + F.nbits + nC);
    }
    uint64_t ax[W];
    uint32_t nb = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      ax[w] = uni64(recs[xpos].a[w]);
      nb += __popcll(ax[w]);
    }
    // speculative child list
    uint32_t ctier = 1;
    uint64_t ctop = 0;
    uint64_t need = (uint64_t)F.len * sizeof(R_t);
    uint64_t crec_addr = arena_take(m, lds_top, hbm_top, need, ctier, ctop);
    if (crec_addr == 0) {
      // no room for the child list: report and abandon the task
      ctr.overflow++;
      if (F.tier == 0) lds_top = F.top; else hbm_top = F.top;
      while (depth > 0) {
        --depth;
        Frame G = m.fr[depth];
        if (uni(G.tier) == 0) lds_top = uni(G.top); else hbm_top = uni(G.top);
      }
      return;
    }
    R_t* crec = (R_t*)crec_addr;
    bool nonmax = false;
    uint64_t absorbed = 0;  // per-lane key sum of entries joining R'
    uint32_t clen = 0, cfirst = kNone, next = kNone;
    for (uint32_t base = 0; base < F.len; base += kWave) {
      uint32_t k = base + lane;
      bool valid = k < F.len;
      uint32_t f = 0;
      uint64_t ae[W], me[W];
      bool sub = true, zero = true, inside = true;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        ae[w] = 0;
        me[w] = 0;
      }
      if (valid) {
        f = recs[k].idx;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          ae[w] = recs[k].a[w];
          me[w] = ae[w] & ax[w];
          sub = sub && (me[w] == ax[w]);
          zero = zero && (me[w] == 0);
          inside = inside && ((ae[w] & ~ax[w]) == 0);
        }
      }
      const bool before = k < xpos, after = k > xpos;
      const bool br = (f & kBranch) != 0;
      const uint32_t e = f & ~kBranch;
      if (valid && sub && before) nonmax = true;
      if (valid && sub && !before) absorbed += m.kC[e];
      const bool pruned = prune && valid && after && br && inside;
      if (pruned) recs[k].idx = e;
      const bool keep = valid && !sub && !zero;
      const bool cbr = br && after;
      uint64_t km = ballot(keep);
      uint32_t pos = clen + lanes_below(km);
      if (keep) {
        crec[pos].idx = e | (cbr ? kBranch : 0u);
#pragma unroll
        for (int w = 0; w < W; ++w) crec[pos].a[w] = me[w];
      }
      if (cfirst == kNone) {
        uint64_t bm = ballot(keep && cbr);
        if (bm) cfirst = clen + __popcll(km & ((1ull << first_lane(bm)) - 1ull));
      }
      clen += __popcll(km);
      if (next == kNone) {
        uint64_t nm = ballot(valid && after && br && !pruned);
        if (nm) next = base + first_lane(nm);
      }
    }
    nonmax = ballot(nonmax) != 0;
    ctr.nodes++;
    F.cur = next;
    if (nonmax) {
      ctr.nonmax++;
      continue;
    }
    // L' key sum: bits of ax over kL
    uint64_t sl = 0;
#pragma unroll
    for (int w = 0; w < W; ++w)
      if ((ax[w] >> lane) & 1ull) sl += m.kL[w * 64 + lane];
    uint64_t sumL = uni64(wave_sum64(sl));
    uint64_t sumR = F.sumR + uni64(wave_sum64(absorbed));
    add_biclique(ctr, biclique_hash(sumL, sumR, P.g.swapped));
    if (P.list) {
      // R' = R0 ∪ {C[i] : a_i ⊇ L'} over the root list
      const R_t* root = (const R_t*)root_rec;
      uint32_t nr = nR0;
      for (uint32_t base = 0; base < nC; base += kWave) {
        uint32_t i = base + lane;
        bool s = false;
        if (i < nC) {
          s = true;
#pragma unroll
          for (int w = 0; w < W; ++w) s = s && ((root[i].a[w] & ax[w]) == ax[w]);
        }
        nr += __popcll(ballot(s));
      }
      uint64_t at = list_reserve(P, 2ull + nb + nr);
      if (at != ~0ull) {
        uint32_t* o = P.list + at;
        if (lane == 0) {
          o[0] = nb;
          o[1] = nr;
        }
        // L' ids: the task L is kept in m.T in list mode
        uint32_t c = 0;
        for (int w = 0; w < W; ++w) {
          bool b = (ax[w] >> lane) & 1ull;
          uint64_t bmk = ballot(b);
          if (b) o[2 + c + lanes_below(bmk)] = m.T[w * 64 + lane];
          c += __popcll(bmk);
        }
        for (uint32_t i = lane; i < nR0; i += kWave) o[2 + nb + i] = m.R[i];
        c = nR0;
        for (uint32_t base = 0; base < nC; base += kWave) {
          uint32_t i = base + lane;
          bool s = false;
          if (i < nC) {
            s = true;
#pragma unroll
            for (int w = 0; w < W; ++w) s = s && ((root[i].a[w] & ax[w]) == ax[w]);
          }
          uint64_t sm = ballot(s);
          if (s) o[2 + nb + c + lanes_below(sm)] = m.C[i];
          c += __popcll(sm);
        }
      }
    }
    if (nb > 1 && cfirst != kNone) {
      // descend: the child list becomes a frame
      if (lane == 0) m.fr[depth] = F;
      wave_mem_wait();
      ++depth;
      if (ctier == 0) lds_top = (uint32_t)(ctop + (uint64_t)clen * sizeof(R_t));
      else hbm_top = ctop + (uint64_t)clen * sizeof(R_t);
      F.sumR = sumR;
      F.rec = crec_addr;
      F.len = clen;
      F.cur = cfirst;
      F.nbits = nb;
      F.tier = ctier;
      F.top = (uint32_t)ctop;
    }
  }
}

// Build the root list of a task (entry i: candidate i, branchable iff
// i >= next_id, bitset of L ∩ N(C[i]) with Wt words; bitsets wider than
// ceil(|L|/64) words are zero-padded) and run the stack machine.
template <int Wt>
__device__ void enumerate_task(const Params& P, const WaveMem& m, const uint32_t* L, uint32_t nL,
                               uint32_t nC, uint32_t next_id, uint64_t sumR, uint32_t nR,
                               Counters& ctr) {
  const DevGraph& g = P.g;
  const uint32_t lane = lane_id();
  int32_t* cmap = m.cmap;
  const uint64_t rs = sizeof(Rec<Wt>);
  uint32_t lds_top = 0;
  uint64_t hbm_top = 0;
  uint32_t rtier = 1;
  uint64_t rtop = 0;
  uint64_t root = arena_take(m, lds_top, hbm_top, (uint64_t)nC * rs, rtier, rtop);
  if (root == 0) {
    ctr.overflow++;
    for (uint32_t i = lane; i < nC; i += kWave) st_wg(cmap + m.C[i], 0);
    wave_mem_wait();
    return;
  }
  if (rtier == 0) lds_top = (uint32_t)(rtop + (uint64_t)nC * rs);
  else hbm_top = rtop + (uint64_t)nC * rs;
  uint64_t* rw = (uint64_t*)root;
  const uint64_t stride = Wt + 1;
  for (uint64_t i = lane; i < (uint64_t)nC * stride; i += kWave) {
    uint64_t e = i / stride, w = i % stride;
    rw[i] = w == 0 ? (uint64_t)((uint32_t)e | (e >= next_id ? kBranch : 0u)) : 0ull;
  }
  wave_mem_wait();
  for_each_edge(L, nL, g.u_off, g.u_adj, m.s_off, m.s_start, [&](uint32_t j, uint32_t w) {
    int32_t p = ld_wg(cmap + w);
    if (p < 0) {
      uint64_t* word = rw + (uint64_t)(-p - 1) * stride + 1 + (j >> 6);
      or_wg(word, 1ull << (j & 63));
    }
  });
  wave_mem_wait();
  for (uint32_t i = lane; i < nC; i += kWave) st_wg(cmap + m.C[i], 0);
  wave_mem_wait();
  dfs<Wt>(P, m, nL, nC, next_id, sumR, root, rtier, rtop, lds_top, hbm_top, ctr, nR);
}

// ---------------------------------------------------------------------------
// one task

template <int WMAX>
__device__ void run_task(const Params& P, const WaveMem& m, const Task& t, int k,
                         Counters& ctr) {
  const DevGraph& g = P.g;
  const uint32_t lane = lane_id();
  ctr.tasks++;
  ctr.nodes++;
  // L = ∩ N(S) (get_common_neighbors, src/IterFinderGpu.cu:680-711)
  const uint32_t* L;
  uint32_t nL, np = 0;
  const uint32_t vk = t.v[k - 1];
  if (k == 1) {
    L = g.v_adj + g.v_off[vk];
    nL = (uint32_t)(g.v_off[vk + 1] - g.v_off[vk]);
  } else {
    // prefix: S[0..k-1) into Lp, then ∩ N(vk) into L
    uint32_t v0 = t.v[0];
    np = (uint32_t)(g.v_off[v0 + 1] - g.v_off[v0]);
    const uint32_t* cur = g.v_adj + g.v_off[v0];
    uint32_t* bufA = m.Lp;
    uint32_t* bufB = m.T;
    for (int i = 1; i < k - 1; ++i) {
      uint32_t vi = t.v[i];
      uint32_t* dst = (i & 1) ? bufB : bufA;
      np = wave_intersect(cur, np, g.v_adj + g.v_off[vi], (uint32_t)(g.v_off[vi + 1] - g.v_off[vi]),
                          dst);
      cur = dst;
    }
    if (cur != m.Lp) {
      for (uint32_t i = lane; i < np; i += kWave) m.Lp[i] = cur[i];
      wave_sync();
    }
    nL = wave_intersect(m.Lp, np, g.v_adj + g.v_off[vk], (uint32_t)(g.v_off[vk + 1] - g.v_off[vk]),
                        m.L);
    L = m.L;
  }
  uint64_t sdeg = 0;  // sum of the prefix vertices' degrees
  for (int i = 0; i < k; ++i) sdeg += (uint64_t)(g.v_off[t.v[i] + 1] - g.v_off[t.v[i]]);
  if (nL == 0) {  // isolated vertex: no biclique (undefined in the reference, Q5)
    ctr.balg += 4ull * sdeg;
    ctr.nonmax++;
    return;
  }
  // 2-hop counts
  int32_t* cmap = m.cmap;
  uint64_t* bm0 = m.bm0;
  uint64_t* bm1 = m.bm1;
  // B(root) = 4 (deg(S) + sum of deg(u), u in L)
  const uint64_t root_edges = for_each_edge(L, nL, g.u_off, g.u_adj, m.s_off, m.s_start,
                                            [&](uint32_t, uint32_t w) {
    int32_t old = add_wg(cmap + w, 1);
    if (old == 0) {
      uint64_t o0 = or_wg(bm0 + (w >> 6), 1ull << (w & 63));
      if (o0 == 0) or_wg(bm1 + (w >> 12), 1ull << ((w >> 6) & 63));
    }
  });
  ctr.balg += 4ull * (sdeg + root_edges);
  wave_mem_wait();
  // sorted extraction of the 2-hop set with its counts; clears the bitmaps
  uint32_t nU2 = 0;
  for (uint32_t b1 = 0; b1 < P.w.nW1; b1 += kWave) {
    uint32_t i1 = b1 + lane;
    uint64_t s = i1 < P.w.nW1 ? ld_wg(bm1 + i1) : 0ull;
    if (s) st_wg(bm1 + i1, 0ull);
    uint64_t nz = ballot(s != 0);
    while (nz) {
      uint32_t src = first_lane(nz);
      nz &= nz - 1;
      uint64_t sw = read_lane64(s, src);
      uint32_t w0 = (b1 + src) * 64 + lane;
      uint64_t word = ((sw >> lane) & 1ull) ? ld_wg(bm0 + w0) : 0ull;
      if (word) st_wg(bm0 + w0, 0ull);
      uint32_t pc = __popcll(word);
      uint32_t incl = wave_incl_scan32(pc);
      uint32_t pos = nU2 + incl - pc;
      while (word) {
        uint32_t b = __builtin_ctzll(word);
        word &= word - 1;
        uint32_t id = w0 * 64 + b;
        m.U[pos] = id;
        m.Ucnt[pos] = (uint32_t)ld_wg(cmap + id);
        ++pos;
      }
      nU2 += read_lane(incl, kWave - 1);
    }
  }
  wave_mem_wait();
  // R = closure, C = the rest (src/IterFinderGpu.cu:794,830-834)
  uint32_t nR = 0, nC = 0, next_id = 0;
  for (uint32_t base = 0; base < nU2; base += kWave) {
    uint32_t i = base + lane;
    bool valid = i < nU2;
    uint32_t w = valid ? m.U[i] : 0, c = valid ? m.Ucnt[i] : 0;
    bool isR = valid && c == nL, isC = valid && c != nL;
    uint64_t mR = ballot(isR), mC = ballot(isC);
    if (isR) m.R[nR + lanes_below(mR)] = w;
    if (isC) m.C[nC + lanes_below(mC)] = w;
    next_id += __popcll(ballot(isC && w < vk));
    nR += __popcll(mR);
    nC += __popcll(mC);
  }
  wave_mem_wait();
  // canonicity
  bool canonical;
  if (k == 1) {
    canonical = nR > 0 && m.R[0] == vk;
  } else {
    canonical = true;
    for (uint32_t i = 0; i < nR && canonical; ++i) {
      uint32_t w = uni(m.R[i]);
      if (w >= vk) break;
      bool in_prefix = false;
      for (int j = 0; j < k - 1; ++j) in_prefix |= (t.v[j] == w);
      if (in_prefix) continue;
      uint32_t c = wave_intersect_count(m.Lp, np, g.v_adj + g.v_off[w],
                                        (uint32_t)(g.v_off[w + 1] - g.v_off[w]));
      canonical = uni(c) == np;
    }
  }
  auto clear_cmap = [&]() {
    for (uint32_t i = lane; i < nU2; i += kWave) st_wg(cmap + m.U[i], 0);
    wave_mem_wait();
  };
  if (!canonical) {
    ctr.nonmax++;
    clear_cmap();
    return;
  }
  // the task's own biclique (L, R); counted once the task is known to stay on
  // this kernel
  auto count_root = [&](uint64_t& sumL, uint64_t& sumR) {
    uint64_t sl = 0, sr = 0;
    for (uint32_t j = lane; j < nL; j += kWave) sl += g.kU[L[j]];
    for (uint32_t i = lane; i < nR; i += kWave) sr += g.kV[m.R[i]];
    sumL = uni64(wave_sum64(sl));
    sumR = uni64(wave_sum64(sr));
    add_biclique(ctr, biclique_hash(sumL, sumR, g.swapped));
    if (P.list) {
      uint64_t at = list_reserve(P, 2ull + nL + nR);
      if (at != ~0ull) {
        uint32_t* o = P.list + at;
        if (lane == 0) {
          o[0] = nL;
          o[1] = nR;
        }
        for (uint32_t j = lane; j < nL; j += kWave) o[2 + j] = L[j];
        for (uint32_t i = lane; i < nR; i += kWave) o[2 + nL + i] = m.R[i];
      }
    }
  };
  uint64_t sumL = 0, sumR = 0;
  uint32_t rest = nC - next_id;
  if (nL <= 1 || rest == 0) {
    count_root(sumL, sumR);
    clear_cmap();
    return;
  }
  const uint32_t W = (nL + 63) / 64;
  // split (src/IterFinderGpu.cu:835-858): children S + [C[i]], i >= next_id
  bool rule = P.split && ((rest > (uint32_t)P.bound_height || nL > (uint32_t)P.bound_height) &&
                          (uint64_t)rest * nL > (uint64_t)P.bound_size);
  bool must = W > (uint32_t)kMaxW;  // |L| beyond any bitset: split regardless
  if ((rule && k < P.max_prefix) || (must && k < kMaxPrefix)) {
    const unsigned long long ne = (rest + kChunk - 1) / kChunk;
    if (lane == 0)
      __hip_atomic_fetch_add(P.s.work_left, ne, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t pushed = 0;
    if (queue_push_children(P.s, t, k, m.C + next_id, rest, pushed)) {
      ctr.splits++;
      count_root(sumL, sumR);
      clear_cmap();
      return;
    }
    if (lane == 0)
      __hip_atomic_fetch_add(P.s.work_left, (unsigned long long)(-(long long)ne),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (must) {
    ctr.overflow++;
    clear_cmap();
    return;
  }
  if (W > (uint32_t)WMAX) {
    // too wide for this kernel: hand the whole task to the wide-bitset pass
    clear_cmap();
    ctr.tasks--;
    ctr.nodes--;
    ctr.balg -= 4ull * (sdeg + root_edges);
    unsigned long long at = ~0ull;
    if (lane == 0) {
      at = __hip_atomic_fetch_add(P.s.n_deferred, 1ull, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
      if (at >= P.s.deferred_cap) at = ~0ull;
    }
    at = uni64(at);
    if (at == ~0ull) {
      ctr.overflow++;
      return;
    }
    if (lane < (uint32_t)kMaxPrefix) st_dev(&P.s.deferred[at].v[lane], t.v[lane]);
    return;
  }
  count_root(sumL, sumR);
  // candidate index map: cmap[C[i]] = -(i+1), cmap[R[i]] = 0
  for (uint32_t i = lane; i < nR; i += kWave) st_wg(cmap + m.R[i], 0);
  for (uint32_t i = lane; i < nC; i += kWave) {
    st_wg(cmap + m.C[i], -(int32_t)(i + 1));
    m.kC[i] = g.kV[m.C[i]];
  }
  for (uint32_t j = lane; j < nL; j += kWave) {
    m.kL[j] = g.kU[L[j]];
    if (P.list) m.T[j] = L[j];
  }
  if (W <= 1) {
    enumerate_task<1>(P, m, L, nL, nC, next_id, sumR, nR, ctr);
  } else if (W <= 2) {
    enumerate_task<2>(P, m, L, nL, nC, next_id, sumR, nR, ctr);
  } else if (W <= 4) {
    enumerate_task<4>(P, m, L, nL, nC, next_id, sumR, nR, ctr);
  } else if constexpr (WMAX >= 16) {
    if (W <= 8) enumerate_task<8>(P, m, L, nL, nC, next_id, sumR, nR, ctr);
    else enumerate_task<16>(P, m, L, nL, nC, next_id, sumR, nR, ctr);
  }
}

}  // namespace mbe
