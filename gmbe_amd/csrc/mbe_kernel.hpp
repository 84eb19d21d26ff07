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
constexpr int kMaxPrefix = 8;              // task capacity (vertices)
constexpr int kShards = 32;                // independent queue rings
constexpr int kCtlStride = 32;             // u64 between control counters (256 B)
constexpr int kMaxW = 16;                  // widest bitset (|L| <= 1024)
constexpr int kMainW = 4;                  // widest bitset of the main kernel
constexpr int kMaxDom = 16;                // dominators tried per list (reduce_list)
constexpr uint32_t kLdsFrames = 32;        // DFS frames kept in LDS (deeper ones in the slab)
// flag in an entry's minidx: the entry is not pushed at this node (its push
// would be non-maximal); children copy minidx without it
constexpr uint32_t kSkip = 0x80000000u;

struct Task {
  uint32_t v[kMaxPrefix];
  uint32_t gid;  // graph of the task (index into the graph table)
};

// One entry of the device graph table (a finder has one, a batch many).
struct DevGraph {
  const int64_t* v_off;
  const uint32_t* v_adj;
  const int64_t* u_off;
  const uint32_t* u_adj;
  const uint64_t* kU;
  const uint64_t* kV;
  uint32_t nU, nV, dmaxV, swapped;
  int32_t bound_height, bound_size;  // split thresholds for this graph
  uint32_t pad[2];
};

// Header word of a queue entry of split children: prefix length k (<= 8),
// child count (<= kChunk) and graph id.  Bit 31 is clear (item entries set it).
__host__ __device__ __forceinline__ uint32_t child_hdr(uint32_t k, uint32_t cnt, uint32_t gid) {
  return k | (cnt << 4) | (gid << 9);
}
__device__ __forceinline__ uint32_t child_hdr_k(uint32_t h) { return h & 0xfu; }
__device__ __forceinline__ uint32_t child_hdr_cnt(uint32_t h) { return (h >> 4) & 0x1fu; }
__device__ __forceinline__ uint32_t child_hdr_gid(uint32_t h) { return (h >> 9) & 0x3fffffu; }
constexpr uint32_t kMaxGraphs = 1u << 22;

// Counter slots of the result block (u64 each).
enum Acc : int {
  A_MAXIMAL = 0, A_NONMAX, A_HSUM, A_HXOR, A_NODES, A_TASKS, A_SPLITS, A_OVERFLOW,
  A_LIST_OVF, A_BALG,
  // profile (filled when Params::profile is set): cycles per phase, work sizes
  A_CYC_WAIT, A_CYC_PREFIX, A_CYC_EXPAND, A_CYC_EXTRACT, A_CYC_BUILD, A_CYC_DFS, A_CYC_SPLIT,
  A_EDGES, A_ENTRIES, A_OVF_ARENA, A_OVF_WIDE, A_OVF_DEFER, A_ENT_W1, A_NODES_W1, A_CYC_TAIL,
  A_CYC_PASS, A_CYC_MERGE, A_CYC_COMPRESS, A_CYC_PUBLISH, A_PUBLISHES, A_MERGES, A_ABANDONED,
  A_LATE, A_CYC_REDUCE, A_DROPPED, A_SKIPPED,
  A_BLIST,  // with count_bytes: bytes of the lists pushes read and write, CSR rows of task roots
  A_COUNT
};

struct Sched {
  unsigned long long* root_next;  // claim counter over top-level vertices
  unsigned long long* work_left;  // outstanding tasks (roots + children)
  uint32_t* q;                    // kShards rings of kEntryWords-word entries
  unsigned long long* qseq;       // per-slot sequence numbers
  unsigned long long* qctl;       // per shard: head, tail (kCtlStride u64 apart)
  unsigned long long* queued;     // entries pushed and not yet popped (a hint)
  unsigned long long qcap;        // slots per shard, power of two
  unsigned long long n_roots;     // top-level vertices of this run (sample)
  const unsigned long long* n_roots_dev;  // if set, the count is read here instead
  unsigned long long* deadline_word;      // if set: the run's absolute deadline (0 = not yet armed)
  const unsigned long long* root_base;    // batch: first root of the j-th graph in order
  const uint32_t* order;                  // batch: graph id of the j-th graph in order
  uint32_t n_graphs;
  uint32_t stride, phase;         // root r -> vertex nV-1-(r*stride+phase) (phase = first rank)
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
      o_dom, o_arena, o_ctx;
  uint64_t arena_bytes, ctx_bytes;  // ctx_bytes: per context slot
  uint32_t ctx_slots;                // context slots per wave
  uint32_t nW0, nW1;
};

struct Params {
  const DevGraph* gt;        // graph table (device memory)
  unsigned long long* per_graph;  // batch: kPerGraph counters per graph, or null
  Sched s;
  Scratch w;
  unsigned long long* acc;   // A_COUNT counters
  uint32_t* list;            // list-mode output (internal ids) or null
  unsigned long long* list_cursor;
  unsigned long long list_cap;
  int prune, max_prefix, split;
  int profile;               // accumulate per-phase cycle counts
  int group;                 // group twin candidates (one-word bitsets)
  int share;                 // split large tasks through shared contexts
  int share_work;            // child entries x pushable entries that make a node shareable
  int share_queue;           // share nodes only while fewer queue entries wait
  int merge_min;             // merge twins of child lists at least this long
  int merge_branch;          // ... with at least this many pushable entries
  int late_share;            // min list length of a frame handed over late (0 = off)
  int dom_k;                 // dominators tried when a list is reduced (0 = no reduction)
  int dom_skip;              // mark pushes a dominator shows non-maximal (reduce_list)
  int recompute;             // a large task with no free context slot: split into recomputed children
  int count_bytes;           // count the algorithmic bytes of every node (alg_bytes)
  int reduce_min;            // reduce only lists at least this long
  uint64_t budget;           // wall-clock ticks a wave may run (0 = no deadline)
  unsigned long long* exit_stamps;  // if set: wall clock at which each block's waves were done
};

// A wave's counter, kept in LDS (the Counters of a wave live in its block's
// shared memory): the first active lane adds with an LDS atomic whose result
// is unused, so an update never waits, and no register holds the counter.
struct LdsCount {
  unsigned long long v;
  __device__ __forceinline__ bool first() const { return lanes_below(ballot(true)) == 0; }
  __device__ __forceinline__ void operator+=(uint64_t x) {
    if (first())
      __hip_atomic_fetch_add(as_lds(&v), (unsigned long long)x, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ __forceinline__ void operator-=(uint64_t x) { *this += (uint64_t)(0 - x); }
  __device__ __forceinline__ void operator^=(uint64_t x) {
    if (first())
      __hip_atomic_fetch_xor(as_lds(&v), (unsigned long long)x, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ __forceinline__ void operator++(int) { *this += 1; }
  __device__ __forceinline__ void operator--(int) { *this += (uint64_t)0 - 1; }
  __device__ __forceinline__ uint64_t get() const {
    return uni64(__hip_atomic_load(as_lds(&v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
  }
  __device__ __forceinline__ void set(uint64_t x) {
    if (first())
      __hip_atomic_store(as_lds(&v), (unsigned long long)x, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
  }
};

// Per-wave counters (one Counters per wave, in LDS; zeroed at kernel start).
struct Counters {
  LdsCount maximal, nonmax, hsum, hxor, nodes, tasks, splits, overflow;
  LdsCount balg;  // algorithmic bytes: 4 B per id of every list a node must read
  LdsCount cyc[7];  // wait, prefix, expand, extract, build, dfs, split
  LdsCount edges, entries, ovf_arena, ovf_wide, ovf_defer;
  LdsCount ent_w1, nodes_w1;  // the part of entries / pushes with one-word bitsets
  LdsCount tail;              // cycles of the final wait for work
  LdsCount c_pass, c_merge, c_compress, c_publish, publishes, merges;
  LdsCount abandoned;         // subtrees left unfinished at the deadline
  LdsCount late;              // frames handed over by late sharing
  LdsCount c_reduce, dropped, skipped;  // list reduction: cycles, entries
  LdsCount blist;  // bytes of the bitset lists pushes read and write, CSR bytes of task roots
};
constexpr int kCounters = (int)(sizeof(Counters) / sizeof(LdsCount));

// Per-graph result counters of a batch (u64 each, kPerGraph per graph).
enum PerGraph : int {
  G_MAXIMAL = 0, G_HSUM, G_HXOR, G_NONMAX, G_NODES, G_OVERFLOW, G_BALG, G_TASKS,
  G_LEFT,  // units of the graph not yet finished (0 at the end: the graph is complete)
  kPerGraph
};

// Count `delta` units of work of graph gid (batch only; lane 0).
__device__ __forceinline__ void graph_work(const Params& P, uint32_t gid, long long delta) {
  if (P.per_graph && lane_id() == 0)
    __hip_atomic_fetch_add(P.per_graph + (uint64_t)gid * kPerGraph + G_LEFT,
                           (unsigned long long)delta, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Move this wave's result counters into graph gid's slots (batch only).
__device__ __forceinline__ void flush_graph(const Params& P, Counters& c, uint32_t gid) {
  wave_mem_wait();
  const uint64_t add[kPerGraph] = {c.maximal.get(), c.hsum.get(), 0, c.nonmax.get(),
                                   c.nodes.get(), c.overflow.get(), c.balg.get(), c.tasks.get()};
  const uint64_t hx = c.hxor.get();
  if (lane_id() == 0) {
    unsigned long long* a = P.per_graph + (uint64_t)gid * kPerGraph;
    for (int i = 0; i < kPerGraph; ++i)
      if (i != G_HXOR && add[i] != 0)
        __hip_atomic_fetch_add(a + i, (unsigned long long)add[i], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    if (hx != 0)
      __hip_atomic_fetch_xor(a + G_HXOR, (unsigned long long)hx, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
  LdsCount* z[] = {&c.maximal, &c.hsum, &c.hxor, &c.nonmax, &c.nodes, &c.overflow, &c.balg,
                   &c.tasks};
  for (LdsCount* q : z) q->set(0);
  wave_mem_wait();
}

// True once this wave's deadline has passed (never without a budget).  The
// constant-rate wall clock is read per wave; no memory is touched.
__device__ __forceinline__ bool past_deadline(const Params& P, uint64_t deadline) {
  return P.budget != 0 && wall_clock64() > deadline;
}

__device__ __forceinline__ uint64_t cycles() { return __builtin_readcyclecounter(); }
// Adds the cycles since `t0` to phase `ph` (when profiling) and restarts t0.
__device__ __forceinline__ void phase(Counters& c, int on, int ph, uint64_t& t0) {
  if (on) {
    uint64_t t = cycles();
    c.cyc[ph] += t - t0;
    t0 = t;
  }
}

struct Frame {
  uint64_t sumR;     // key sum of R at this node
  uint64_t rec;      // generic address of the first list entry
  uint32_t len;      // entries
  uint32_t cur;      // position of the next entry to push (kNone: exhausted)
  uint32_t curidx;   // its candidate index
  uint32_t nbits;    // |L| at this node
  uint32_t tier;     // 0 = LDS arena, 1 = HBM arena
  uint32_t top;      // arena offset of this list (freed on pop)
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
__device__ __forceinline__ uint32_t wave_intersect(const uint32_t* A, uint32_t nA, const uint32_t* B,
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

__device__ __forceinline__ uint32_t wave_intersect_count(const uint32_t* A, uint32_t nA, const uint32_t* B,
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
__device__ __forceinline__ uint64_t for_each_edge(const uint32_t* L, uint32_t nL, const int64_t* off,
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

__device__ __forceinline__ unsigned long long* shard_head(const Sched& s, uint32_t sh) {
  return s.qctl + (uint64_t)(2 * sh) * kCtlStride;
}
__device__ __forceinline__ unsigned long long* shard_tail(const Sched& s, uint32_t sh) {
  return s.qctl + (uint64_t)(2 * sh + 1) * kCtlStride;
}
__device__ __forceinline__ uint64_t slot_of(const Sched& s, uint32_t sh, unsigned long long pos) {
  return (uint64_t)sh * s.qcap + (pos & (s.qcap - 1));
}

// Pop one entry into `word` (lane l < 32 holds word l), scanning the shards
// from `home`.
__device__ __forceinline__ bool queue_pop(const Sched& s, uint32_t& word, uint32_t home) {
  const uint32_t lane = lane_id();
  if ((long long)uni64(ld_dev(s.queued)) <= 0) return false;
  unsigned long long got = ~0ull;
  uint32_t gsh = 0;
  if (lane == 0) {
    for (uint32_t i = 0; i < (uint32_t)kShards && got == ~0ull; ++i) {
      const uint32_t sh = (home + i) & (kShards - 1);
      unsigned long long h = ld_dev(shard_head(s, sh));
      for (int tries = 0; tries < 4; ++tries) {
        unsigned long long tl = ld_dev(shard_tail(s, sh));
        if (h >= tl) break;
        if (__hip_atomic_compare_exchange_strong(shard_head(s, sh), &h, h + 1, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          got = h;
          gsh = sh;
          break;
        }
      }
    }
    if (got != ~0ull) {
      __hip_atomic_fetch_add(s.queued, (unsigned long long)-1ll, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      const uint64_t slot = slot_of(s, gsh, got);
      while (ld_dev(s.qseq + slot) != got + 1) __builtin_amdgcn_s_sleep(1);
    }
  }
  got = uni64(got);
  if (got == ~0ull) return false;
  gsh = uni(gsh);
  wave_drain();
  const uint64_t slot = slot_of(s, gsh, got);
  word = lane < (uint32_t)kEntryWords ? ld_dev(s.q + slot * kEntryWords + lane) : 0u;
  wave_drain();
  if (lane == 0) st_dev(s.qseq + slot, got + s.qcap);
  return true;
}

// Reserve ne consecutive positions in one shard, trying `home` first.
// Returns false when every shard lacks room.
__device__ __forceinline__ bool queue_reserve(const Sched& s, uint32_t home, uint32_t ne, uint32_t& sh_out,
                              unsigned long long& start) {
  const uint32_t lane = lane_id();
  unsigned long long st = ~0ull;
  uint32_t sh_got = 0;
  if (lane == 0) {
    for (uint32_t i = 0; i < (uint32_t)kShards && st == ~0ull; ++i) {
      const uint32_t sh = (home + i) & (kShards - 1);
      unsigned long long tl = ld_dev(shard_tail(s, sh));
      for (;;) {
        unsigned long long hd = ld_dev(shard_head(s, sh));
        if (tl + ne > hd + s.qcap) break;
        if (__hip_atomic_compare_exchange_strong(shard_tail(s, sh), &tl, tl + ne,
                                                 __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
          st = tl;
          sh_got = sh;
          break;
        }
      }
    }
  }
  start = uni64(st);
  sh_out = uni(sh_got);
  return start != ~0ull;
}

// Push the children prefix+[cand[i]], i < n, as ceil(n / kChunk) entries.
// All or nothing: false (nothing pushed) when the queue lacks room.
__device__ __forceinline__ bool queue_push_children(const Sched& s, const Task& prefix, int k,
                                    const uint32_t* cand, uint32_t n, uint32_t& n_entries,
                                    uint32_t home) {
  uint32_t lane = lane_id();
  const uint32_t ne = (n + kChunk - 1) / kChunk;
  n_entries = ne;
  uint32_t sh = 0;
  unsigned long long start = 0;
  if (!queue_reserve(s, home, ne, sh, start)) return false;
  // two entries per pass: lanes 0-31 and 32-63
  for (uint32_t e0 = 0; e0 < ne; e0 += 2) {
    uint32_t e = e0 + (lane >> 5), w = lane & 31;
    if (e < ne && w == 0) {
      const unsigned long long pos = start + e;
      while (ld_dev(s.qseq + slot_of(s, sh, pos)) != pos) __builtin_amdgcn_s_sleep(1);
    }
    wave_drain();
    if (e < ne) {
      const uint64_t slot = slot_of(s, sh, start + e);
      uint32_t cnt = n - e * kChunk < (uint32_t)kChunk ? n - e * kChunk : (uint32_t)kChunk;
      uint32_t val;
      if (w == 0) val = child_hdr((uint32_t)k, cnt, prefix.gid);
      else if (w <= 7) val = w - 1 < (uint32_t)k ? prefix.v[w - 1] : kNone;
      else val = (w - 8) < cnt ? cand[e * kChunk + (w - 8)] : kNone;
      st_dev(s.q + slot * kEntryWords + w, val);
    }
    wave_drain();  // the words are in place before the slot is published
    if (e < ne && w == 0) st_dev(s.qseq + slot_of(s, sh, start + e), start + e + 1);
  }
  wave_drain();
  if (lane == 0)
    __hip_atomic_fetch_add(s.queued, (unsigned long long)ne, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// Top-level vertices of this launch (read from memory when a previous launch
// decided them, as for the wide pass of a captured batch).
__device__ __forceinline__ unsigned long long roots_total(const Sched& s) {
  return s.n_roots_dev ? uni64(ld_dev(s.n_roots_dev)) : s.n_roots;
}

// Next unit of work: a queue entry (split children) first, then a top-level
// vertex in descending id (src/IterFinderGpu.cu:992-1018).  Returns 0 once
// every unit has finished (the gc/lc/tc protocol of
// src/IterFinderGpu.cu:978-986,1020-1034 collapsed into one outstanding-work
// counter), 1 for an entry (in `word`), 2 for a root task (in `t`, `k`).
__device__ __forceinline__ int next_work(const Params& P, uint32_t& word, Task& t, int& k, uint32_t home,
                         uint64_t deadline) {
  const Sched& s = P.s;
  uint32_t lane = lane_id();
  uint32_t nap = 1;
  const unsigned long long n_roots = roots_total(s);
  for (;;) {
    if (past_deadline(P, deadline)) return 0;  // leave what is queued: the run is cut short
    if (queue_pop(s, word, home)) return 1;
    unsigned long long r = ~0ull;
    if (lane == 0 && ld_dev(s.root_next) < n_roots) {
      r = __hip_atomic_fetch_add(s.root_next, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    r = uni64(r);
    if (r != ~0ull && r < n_roots) {
      if (s.root_tasks) {
        uint32_t w = lane < (uint32_t)kMaxPrefix ? s.root_tasks[r].v[lane] : kNone;
        k = 0;
        for (int i = 0; i < kMaxPrefix; ++i) {
          t.v[i] = read_lane(w, i);
          k += t.v[i] != kNone;
        }
        t.gid = uni(s.root_tasks[r].gid);
        return 2;
      }
      uint32_t gid = 0;
      uint64_t rank = r * s.stride + s.phase;
      if (s.root_base) {
        // batch: the graph whose root range holds r (graphs in processing order)
        uint32_t lo = 0, hi = s.n_graphs;  // root_base[lo] <= r < root_base[hi]
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (s.root_base[mid] <= r) lo = mid; else hi = mid;
        }
        lo = uni(lo);
        gid = uni(s.order[lo]);
        rank = r - s.root_base[lo];
      }
      for (int i = 0; i < kMaxPrefix; ++i) t.v[i] = kNone;
      t.v[0] = P.gt[gid].nV - 1 - (uint32_t)rank;
      t.gid = gid;
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
  Frame* fr;       // DFS frames (slab; the first kLdsFrames live in fr_lds)
  Frame* fr_lds;   // DFS frames in LDS
  uint64_t* dom;   // dominator bitsets of reduce_list (kMaxDom x kMaxW)
  char* arena;     // HBM list arena
  uint64_t arena_bytes;
  char* lds;       // LDS list arena
  uint32_t lds_bytes;
  uint32_t* s_off; // 64 + 1 u32 in LDS
  int64_t* s_start;// 64 i64 in LDS
  const DevGraph* g;  // graph of the current unit of work (set per unit)
  uint64_t deadline;  // wall clock past which this wave stops (with a budget)
};

// Whether p lies in this wave's LDS list arena.
__device__ __forceinline__ bool in_lds(const WaveMem& m, const void* p) {
  return (uint64_t)((const char*)p - m.lds) < (uint64_t)m.lds_bytes;
}

// Frame i of the wave's DFS stack: LDS for the shallow ones.
__device__ __forceinline__ void frame_put(const WaveMem& m, uint32_t i, const Frame& F) {
  if (lane_id() == 0) {
    if (i < kLdsFrames) {
      auto* q = as_lds(m.fr_lds) + i;
      q->sumR = F.sumR;
      q->rec = F.rec;
      q->len = F.len;
      q->cur = F.cur;
      q->curidx = F.curidx;
      q->nbits = F.nbits;
      q->tier = F.tier;
      q->top = F.top;
    } else {
      m.fr[i] = F;
    }
  }
  wave_mem_wait();
}
__device__ __forceinline__ Frame frame_get(const WaveMem& m, uint32_t i) {
  Frame F;
  if (i < kLdsFrames) {
    const auto* q = as_lds(m.fr_lds) + i;
    F.sumR = q->sumR;
    F.rec = q->rec;
    F.len = q->len;
    F.cur = q->cur;
    F.curidx = q->curidx;
    F.nbits = q->nbits;
    F.tier = q->tier;
    F.top = q->top;
  } else {
    F = m.fr[i];
  }
  F.sumR = uni64(F.sumR);
  F.rec = uni64(F.rec);
  F.len = uni(F.len);
  F.cur = uni(F.cur);
  F.curidx = uni(F.curidx);
  F.nbits = uni(F.nbits);
  F.tier = uni(F.tier);
  F.top = uni(F.top);
  return F;
}

__device__ __forceinline__ uint32_t graph_id(const Params& P, const WaveMem& m) {
  return (uint32_t)(m.g - P.gt);
}

__device__ __forceinline__ WaveMem wave_mem(const Params& P, uint64_t gw, char* lds,
                                            uint32_t lds_bytes, uint32_t* s_off,
                                            int64_t* s_start, Frame* fr_lds) {
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
  m.fr_lds = fr_lds;
  m.dom = (uint64_t*)(b + S.o_dom);
  m.arena = b + S.o_arena;
  m.arena_bytes = S.arena_bytes;
  m.lds = lds;
  m.lds_bytes = lds_bytes;
  m.s_off = s_off;
  m.s_start = s_start;
  m.g = P.gt;
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
__device__ __forceinline__ uint64_t list_reserve(const Params& P, uint64_t words) {
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
// the stack machine over one task (W = bitset words per entry)
//
// A list entry is a group of candidates whose neighbourhoods inside the
// node's L are identical.  Such twins behave identically under every later
// push: the first one that is pushed absorbs the others into R' and PRUNE_EN
// then prunes them, and a twin before it makes it non-canonical.  So a group
// is fully described by
//   minidx  its smallest member index,
//   rep     the member that may still be pushed (kNone when none may),
//   ksum    the members' hash-key sum,
//   a       the bitset,
// and one entry per group gives the reference's pushes, verdicts and prunes.
// Entries that become twins inside a child's smaller L are merged again
// (minidx and rep by minimum, ksum by sum), so a node's list never holds more
// than 2^|L| - 2 entries.  Entries of a list are in no particular order:
// pushes go by increasing rep.

template <int W>
struct Rec {
  uint32_t minidx;
  uint32_t rep;
  uint64_t ksum;  // hash-key sum of the group's members
  uint64_t a[W];
};

// Allocate `bytes` for a list: LDS first, then HBM.  Returns the generic
// address or 0 when neither arena has room.
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

__device__ __forceinline__ uint64_t wave_min64(uint64_t x) {
  for (int s = 32; s > 0; s >>= 1) {
    uint32_t lo = shfl_xor32((uint32_t)x, s), hi = shfl_xor32((uint32_t)(x >> 32), s);
    uint64_t y = ((uint64_t)hi << 32) | lo;
    x = y < x ? y : x;
  }
  return x;
}

// Task-level arrays a stack machine reads: hash keys of L (by L index) and
// candidate vertex ids.  When REMOTE, they belong to another wave's shared
// context and are read with device-scope loads.
struct TaskArrays {
  const uint64_t* kL;
  const uint32_t* C;
};

template <bool REMOTE, class T>
__device__ __forceinline__ T rd(const T* p) {
  if constexpr (REMOTE) return ld_dev(p);
  else return *p;
}

struct PassOut {
  uint32_t cbranch;     // pushable entries of the child list
  bool nonmax;
  uint64_t absorbed;    // key sum of the groups joining R'
  uint64_t best_next;   // (rep << 32 | position) of the next push in the parent
  uint64_t best_child;  // same, for the child's first push
  uint32_t clen;
};

// One push: x (candidate index xidx, bitset ax) against the parent's list
// `recs`, writing the child list to `crec`.  LIST_REMOTE: the parent list is
// another wave's shared context (read-only: no prune writes).
// ordered: the list is in ascending minidx and every pushable entry has
// rep == minidx (reduce_list and merge_twins keep this when dom_skip is on),
// so pushes go in position order and the next push of the parent, and the
// first of the child, are the first qualifying positions (a ballot per
// chunk); otherwise they are minimum reps over the wave.
// LDS_IN / LDS_OUT: the parent / child list is in LDS (ds_ accesses).
template <int W, bool LIST_REMOTE, bool LDS_IN = false, bool LDS_OUT = false>
__device__ __forceinline__ PassOut push_pass(Rec<W>* recs, uint32_t len, const uint64_t* ax, uint32_t xidx,
                             Rec<W>* crec, bool prune, bool ordered) {
  const uint32_t lane = lane_id();
  bool nonmax = false;
  // once the node is known to be non-maximal its child list is never used:
  // stop writing it (the prune marks and the next push still need every entry)
  bool dead = false;
  uint64_t absorbed = 0, best_next = ~0ull, best_child = ~0ull;
  uint32_t clen = 0, cbranch = 0;
  for (uint32_t base = 0; base < len; base += kWave) {
    const uint32_t k = base + lane;
    const bool valid = k < len;
    uint32_t mi = kNone, rep = kNone;
    uint64_t ks = 0, ae[W];
#pragma unroll
    for (int w = 0; w < W; ++w) ae[w] = 0;
    if (valid) {
      if constexpr (LDS_IN) {
        const auto* q = as_lds(recs) + k;
        mi = q->minidx;
        rep = q->rep;
        ks = q->ksum;
#pragma unroll
        for (int w = 0; w < W; ++w) ae[w] = q->a[w];
      } else {
        mi = rd<LIST_REMOTE>(&recs[k].minidx);
        rep = rd<LIST_REMOTE>(&recs[k].rep);
        ks = rd<LIST_REMOTE>(&recs[k].ksum);
#pragma unroll
        for (int w = 0; w < W; ++w) ae[w] = rd<LIST_REMOTE>(&recs[k].a[w]);
      }
    }
    // maximality first, for the whole chunk: a chunk that shows the node is
    // non-maximal writes none of its child entries
    bool sub = valid, zero = true, inside = valid;
    uint64_t me[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      me[w] = ae[w] & ax[w];
      sub = sub && (me[w] == ax[w]);
      zero = zero && (me[w] == 0);
      inside = inside && ((ae[w] & ~ax[w]) == 0);
    }
    const uint32_t mim = mi & ~kSkip;
    nonmax = nonmax || (sub && mim < xidx);  // a twin or superset before x
    // groups that join R': few lanes, summed one by one
    uint64_t am = ballot(sub && mim >= xidx);
    while (am) {
      const uint32_t l = (uint32_t)__builtin_ctzll(am);
      am &= am - 1;
      absorbed += read_lane64(ks, l);
    }
    dead = dead || ballot(nonmax) != 0;
    const bool br = rep != kNone && rep > xidx;
    const bool pruned = prune && valid && br && inside;
    if constexpr (!LIST_REMOTE) {
      if (pruned) {
        if constexpr (LDS_IN) as_lds(recs)[k].rep = kNone;
        else recs[k].rep = kNone;
      }
    }
    const bool keep = !dead && valid && !sub && !zero;
    const uint32_t crep = br ? rep : kNone;  // branchability before this push's prune
    const uint64_t km = ballot(keep);
    const uint32_t pos = clen + lanes_below(km);
    if (keep) {
      if constexpr (LDS_OUT) {
        auto* q = as_lds(crec) + pos;
        q->minidx = mim;  // a skip mark holds at its own node only
        q->rep = crep;
        q->ksum = ks;
#pragma unroll
        for (int w = 0; w < W; ++w) q->a[w] = me[w];
      } else {
        crec[pos].minidx = mim;
        crec[pos].rep = crep;
        crec[pos].ksum = ks;
#pragma unroll
        for (int w = 0; w < W; ++w) crec[pos].a[w] = me[w];
      }
    }
    const bool nx = valid && br && !pruned && !(mi & kSkip);
    const bool cc = keep && crep != kNone;
    if (ordered) {
      if (best_next == ~0ull) {
        const uint64_t bm = ballot(nx);
        if (bm) {
          const uint32_t l = (uint32_t)__builtin_ctzll(bm);
          best_next = ((uint64_t)read_lane(rep, l) << 32) | (base + l);
        }
      }
      if (best_child == ~0ull) {
        const uint64_t bm = ballot(cc);
        if (bm) {
          const uint32_t l = (uint32_t)__builtin_ctzll(bm);
          best_child = ((uint64_t)read_lane(crep, l) << 32) | read_lane(pos, l);
        }
      }
    } else {
      if (nx) {
        const uint64_t c = ((uint64_t)rep << 32) | k;
        best_next = c < best_next ? c : best_next;
      }
      if (cc) {
        const uint64_t c = ((uint64_t)crep << 32) | pos;
        best_child = c < best_child ? c : best_child;
      }
    }
    clen += __popcll(km);
    cbranch += __popcll(ballot(cc));
  }
  PassOut o;
  o.cbranch = cbranch;
  o.nonmax = dead;
  o.absorbed = absorbed;
  o.best_next = ordered ? best_next : wave_min_pair(best_next);
  o.best_child = ordered ? best_child : wave_min_pair(best_child);
  o.clen = clen;
  return o;
}

template <int W, bool KL_R>
__device__ __forceinline__ uint64_t key_sum_of_bits(const WaveMem& m, const uint64_t* ax,
                                                    const TaskArrays& ta) {
  const uint32_t lane = lane_id();
  uint64_t sl = 0;
  const bool lds = !KL_R && in_lds(m, ta.kL);
#pragma unroll
  for (int w = 0; w < W; ++w)
    if ((ax[w] >> lane) & 1ull)
      sl += lds ? as_lds(ta.kL)[w * 64 + lane] : rd<KL_R>(ta.kL + w * 64 + lane);
  return wave_sum_u64(sl);
}

// Re-index the bitsets of a list onto the set bits of ax (at most 64): entry
// bit t of the output is the input's bit at the t-th set position of ax.
template <int W>
__device__ __forceinline__ void compress_list(const Rec<W>* src, uint32_t len, const uint64_t* ax, Rec<1>* dst) {
  const uint32_t lane = lane_id();
  lds_flat_fence();  // src / dst may be LDS lists reached by ds_ elsewhere
  for (uint32_t i = lane; i < len; i += kWave) {
    uint64_t out = 0;
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      uint64_t mw = ax[w];
      const uint64_t aw = src[i].a[w];
      while (mw) {
        const uint32_t b = __builtin_ctzll(mw);
        mw &= mw - 1;
        out |= ((aw >> b) & 1ull) << t;
        ++t;
      }
    }
    dst[i].minidx = src[i].minidx & ~kSkip;
    dst[i].rep = src[i].rep;
    dst[i].ksum = src[i].ksum;
    dst[i].a[0] = out;
  }
  lds_flat_fence();
}

// ---------------------------------------------------------------------------
// shared task contexts
//
// A task the split rule marks as large publishes its root list once, in the
// context area of its wave's slab, and pushes one work item (entry position,
// candidate) per pushable group.  Any wave runs an item: one push pass over
// the shared list (device-scope loads), then the stack machine on the child
// list in its own arenas.  This replaces the reference's re-computation of
// every child in LargeToTiny (src/IterFinderGpu.cu:889-910) and in the tiny
// tasks, whose cost here was dominated by re-reading hub neighbourhoods.

struct CtxHdr {
  uint64_t sumR0;           // key sum of R at the shared node
  unsigned long long refs;  // items not yet finished
  uint32_t nL, nC, len, W;  // task |L| and |C|, list length, bitset words
  uint32_t nbits;           // |L| at the shared node
  uint32_t gid;             // graph of the shared node
  uint32_t pad[8];
};

// Context id = wave * ctx_slots + slot; a wave publishes only into its own
// slots, and a slot is free again once its refs count has dropped to zero.
__device__ __forceinline__ CtxHdr* ctx_hdr(const Params& P, uint64_t id) {
  const uint64_t wave = id / P.w.ctx_slots, slot = id % P.w.ctx_slots;
  return (CtxHdr*)(P.w.base + wave * P.w.slab + P.w.o_ctx + slot * P.w.ctx_bytes);
}

// A free context slot of wave gw (its id), or kNone.  Lane s polls slot s.
__device__ __forceinline__ uint32_t free_ctx(const Params& P, uint64_t gw) {
  const uint32_t lane = lane_id();
  const uint64_t id = gw * P.w.ctx_slots + lane;
  bool free_slot = false;
  if (lane < P.w.ctx_slots) free_slot = ld_dev(&ctx_hdr(P, id)->refs) == 0ull;
  const uint64_t b = ballot(free_slot);
  return b ? (uint32_t)(gw * P.w.ctx_slots) + (uint32_t)__builtin_ctzll(b) : kNone;
}
// context layout after the header: list (len recs), kL (nL), C (nC)
template <int W>
__device__ __forceinline__ void ctx_arrays(CtxHdr* h, uint32_t nL, uint32_t nC, uint32_t len,
                                           Rec<W>*& recs, TaskArrays& ta) {
  char* p = (char*)h + sizeof(CtxHdr);
  recs = (Rec<W>*)p;
  p += (uint64_t)len * sizeof(Rec<W>);
  ta.kL = (const uint64_t*)p;
  p += (uint64_t)nL * 8;
  ta.C = (const uint32_t*)p;
}
template <int W>
__device__ __forceinline__ uint64_t ctx_bytes(uint32_t nL, uint32_t nC, uint32_t len) {
  return sizeof(CtxHdr) + (uint64_t)len * sizeof(Rec<W>) + (uint64_t)nC * 4 + (uint64_t)nL * 8;
}

constexpr uint32_t kItemKind = 0x80000000u;
constexpr int kItemsPerEntry = 15;  // (position, candidate) pairs after hdr, owner

// Publish a node's list `root` (len entries, in this wave's arenas) as this
// wave's context and push one item per pushable entry.  The task arrays
// `src` (local or remote) are copied so that the context is self-contained.
// Only entries whose candidate is at least min_rep become items: a node that
// has already pushed its first candidates publishes the rest.  Returns false
// (nothing published) when the queue lacks room or fewer than min_items
// entries are left to push.
template <int W, bool KL_R, bool C_R>
__device__ __forceinline__ bool publish_list(const Params& P, uint64_t gw, uint32_t ctx, uint32_t gid,
                             const Rec<W>* root, uint32_t len,
                             uint32_t nL, uint32_t nC, uint32_t nbits, uint64_t sumR0,
                             const TaskArrays& src, uint32_t min_rep = 0,
                             uint32_t min_items = 1) {
  const uint32_t lane = lane_id();
  const Sched& s = P.s;
  CtxHdr* h = ctx_hdr(P, ctx);
  Rec<W>* recs;
  TaskArrays ta;
  ctx_arrays<W>(h, nL, nC, len, recs, ta);
  lds_flat_fence();  // `root` and src.kL may be in LDS, written with ds_ stores
  // items: pushable groups, in the order of the list
  uint32_t n_items = 0;
  for (uint32_t base = 0; base < len; base += kWave) {
    uint32_t i = base + lane;
    bool it = false;
    if (i < len) {
      const uint32_t r = root[i].rep;
      it = r != kNone && r >= min_rep && !(root[i].minidx & kSkip);
    }
    n_items += __popcll(ballot(it));
  }
  if (n_items == 0 || n_items < min_items) return false;
  const uint32_t ne = (n_items + kItemsPerEntry - 1) / kItemsPerEntry;
  // copy the context with device-scope stores
  for (uint32_t i = lane; i < len; i += kWave) {
    st_dev(&recs[i].minidx, root[i].minidx & ~kSkip);
    st_dev(&recs[i].rep, root[i].rep);
    st_dev(&recs[i].ksum, root[i].ksum);
#pragma unroll
    for (int w = 0; w < W; ++w) st_dev(&recs[i].a[w], root[i].a[w]);
  }
  // candidate ids: read below only to count algorithmic bytes
  if (P.count_bytes)
    for (uint32_t i = lane; i < nC; i += kWave)
      st_dev((uint32_t*)ta.C + i, rd<C_R>(src.C + i));
  for (uint32_t j = lane; j < nL; j += kWave)
    st_dev((uint64_t*)ta.kL + j, rd<KL_R>(src.kL + j));
  if (lane == 0) {
    st_dev(&h->sumR0, sumR0);
    st_dev(&h->nbits, nbits);
    st_dev(&h->nL, nL);
    st_dev(&h->nC, nC);
    st_dev(&h->len, len);
    st_dev(&h->W, (uint32_t)W);
    st_dev(&h->gid, gid);
    st_dev(&h->refs, (unsigned long long)n_items);
  }
  wave_drain();
  // reserve ne entries in one shard
  uint32_t sh = 0;
  unsigned long long start = 0;
  if (!queue_reserve(s, (uint32_t)(gw & (kShards - 1)), ne, sh, start)) {
    if (lane == 0) st_dev(&h->refs, 0ull);
    return false;
  }
  if (lane == 0)
    __hip_atomic_fetch_add(s.work_left, (unsigned long long)ne, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  graph_work(P, gid, (long long)ne);
  // fill the entries: item j goes to entry j / 15, slot j % 15
  uint32_t j0 = 0;
  for (uint32_t base = 0; base < len; base += kWave) {
    uint32_t i = base + lane;
    bool it = false;
    if (i < len) {
      const uint32_t r = root[i].rep;
      it = r != kNone && r >= min_rep && !(root[i].minidx & kSkip);
    }
    uint64_t bm = ballot(it);
    if (it) {
      uint32_t j = j0 + lanes_below(bm);
      const unsigned long long pos = start + j / kItemsPerEntry;
      const uint64_t slot = slot_of(s, sh, pos);
      uint32_t* e = s.q + slot * kEntryWords;
      if (j % kItemsPerEntry == 0) {
        // first item of an entry: wait until the slot is free, write the header
        while (ld_dev(s.qseq + slot) != pos) __builtin_amdgcn_s_sleep(1);
        uint32_t cnt = n_items - j < (uint32_t)kItemsPerEntry ? n_items - j : kItemsPerEntry;
        st_dev(e + 0, kItemKind | cnt);
        st_dev(e + 1, ctx);
      }
    }
    wave_drain();
    if (it) {
      uint32_t j = j0 + lanes_below(bm);
      const uint64_t slot = slot_of(s, sh, start + j / kItemsPerEntry);
      uint32_t* e = s.q + slot * kEntryWords;
      st_dev(e + 2 + 2 * (j % kItemsPerEntry), i);
      st_dev(e + 3 + 2 * (j % kItemsPerEntry), root[i].rep);
    }
    j0 += __popcll(bm);
  }
  wave_drain();
  for (uint32_t e = lane; e < ne; e += kWave) st_dev(s.qseq + slot_of(s, sh, start + e), start + e + 1);
  wave_drain();
  if (lane == 0)
    __hip_atomic_fetch_add(s.queued, (unsigned long long)ne, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// Merge the entries of a list (len entries at crec) whose bitsets are
// identical: minidx and rep by minimum, ksum by sum, in place.  Open
// addressing in a table taken from the arenas (LDS first); a slot's identity
// word packs a 32-bit hash of the bitset with the index of its first-inserted
// entry, set by one CAS, and a probe that meets the same hash compares the
// full bitsets, so no lane ever waits for another.  Each entry then parks its
// slot number in its minidx (already folded into the slot), and a stable
// compaction keeps the smallest entry of every slot with the merged fields:
// it writes only at or below the position it reads, and never reads another
// entry.  Returns the new length and recomputes the first push and the
// pushable count; leaves the list unchanged (returns len) when the arenas
// lack room for the table.
template <int W>
__device__ __forceinline__ uint32_t merge_twins_impl(const WaveMem& m, Rec<W>* crec, uint32_t len,
                                                     uint32_t& lds_top, uint64_t& hbm_top,
                                                     uint64_t& best_child, uint32_t& cbranch,
                                                     bool ordered) {
  const uint32_t lane = lane_id();
  struct Slot {
    unsigned long long id;  // hash << 32 | (first-inserted entry + 1); 0 = empty
    uint32_t mi, rep;
    uint64_t ksum;
    uint32_t first;         // smallest entry of the slot
    uint32_t pad;
  };
  uint32_t T = 64;
  while (T < 2 * len) T <<= 1;
  uint32_t tier = 1;
  uint64_t tb = 0;
  const uint64_t taddr = arena_take(m, lds_top, hbm_top, (uint64_t)T * sizeof(Slot), tier, tb);
  if (taddr == 0) return len;
  Slot* tab = (Slot*)taddr;
  for (uint32_t i = lane; i < T; i += kWave) {
    tab[i].id = 0;
    tab[i].mi = kNone;
    tab[i].rep = kNone;
    tab[i].ksum = 0;
    tab[i].first = kNone;
  }
  wave_mem_wait();
  for (uint32_t i = lane; i < len; i += kWave) {
    uint64_t key[W];
    uint64_t hk = 0x2545F4914F6CDD1Dull;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      key[w] = crec[i].a[w];
      hk = mix64(hk ^ key[w]);
    }
    const unsigned long long me = ((hk >> 32) << 32) | (unsigned long long)(i + 1);
    uint32_t h = (uint32_t)hk & (T - 1);
    for (;;) {
      unsigned long long old = 0;
      if (__hip_atomic_compare_exchange_strong(&tab[h].id, &old, me, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
        break;  // first entry of this bitset
      if ((old >> 32) == (hk >> 32)) {
        const uint32_t f = (uint32_t)(old & 0xffffffffu) - 1u;
        bool same = true;
#pragma unroll
        for (int w = 0; w < W; ++w) same = same && (crec[f].a[w] == key[w]);
        if (same) break;
      }
      h = (h + 1) & (T - 1);
    }
    __hip_atomic_fetch_min(&tab[h].mi, crec[i].minidx & ~kSkip, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    const uint32_t rp = crec[i].rep;
    if (rp != kNone)
      __hip_atomic_fetch_min(&tab[h].rep, rp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    add_wg(&tab[h].ksum, crec[i].ksum);
    __hip_atomic_fetch_min(&tab[h].first, i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    // every probe that compares against this entry's bitset has its answer
    // from the CAS above; the bitset itself is left untouched
    st_wg(&crec[i].minidx, h);
  }
  wave_mem_wait();
  uint32_t n = 0, nbr = 0;
  uint64_t best = ~0ull;
  for (uint32_t base = 0; base < len; base += kWave) {
    const uint32_t i = base + lane;
    bool keep = false;
    uint32_t h = 0;
    uint64_t a[W];
#pragma unroll
    for (int w = 0; w < W; ++w) a[w] = 0;
    if (i < len) {
      h = ld_wg(&crec[i].minidx);
      keep = ld_wg(&tab[h].first) == i;
#pragma unroll
      for (int w = 0; w < W; ++w) a[w] = crec[i].a[w];
    }
    const uint64_t km = ballot(keep);
    const uint32_t pos = n + lanes_below(km);
    uint32_t rp = kNone;
    wave_mem_wait();  // every lane has read its entry before any is overwritten
    if (keep) {
      rp = ld_wg(&tab[h].rep);
      const uint32_t mi = ld_wg(&tab[h].mi);
      // ordered lists: a group whose smallest member is not its pushable one
      // is non-maximal whenever pushed (that member contains it and comes
      // first), here and below; it becomes excluded, so that every pushable
      // entry keeps rep == minidx
      if (ordered && rp != kNone && rp != mi) rp = kNone;
      crec[pos].minidx = mi;
      crec[pos].rep = rp;
      crec[pos].ksum = ld_wg(&tab[h].ksum);
#pragma unroll
      for (int w = 0; w < W; ++w) crec[pos].a[w] = a[w];
      if (rp != kNone) {
        const uint64_t c = ((uint64_t)rp << 32) | pos;
        best = c < best ? c : best;
      }
    }
    nbr += __popcll(ballot(rp != kNone));
    n += __popcll(km);
  }
  wave_mem_wait();
  best_child = (wave_min_pair(best));
  cbranch = nbr;
  return n;
}

// The list and the table are reached through generic pointers (flat_
// accesses) while an LDS list is read and written with ds_ accesses
// elsewhere: fenced on both sides.
template <int W>
__device__ __forceinline__ uint32_t merge_twins(const WaveMem& m, Rec<W>* crec, uint32_t len,
                                                uint32_t& lds_top, uint64_t& hbm_top,
                                                uint64_t& best_child, uint32_t& cbranch,
                                                bool ordered) {
  lds_flat_fence();
  const uint32_t n = merge_twins_impl<W>(m, crec, len, lds_top, hbm_top, best_child, cbranch,
                                         ordered);
  lds_flat_fence();
  return n;
}

// Reduce a node's list before its first push.  With minrep the smallest
// pushable candidate of the list, an entry with no pushable member and
// minidx < minrep (an excluded entry) precedes every push made anywhere in
// the node's subtree and is never pushed or absorbed there: its only effect
// is the non-maximal verdict of a push whose L' lies inside its bitset
// (finder_push's maximality test, src/IterFinderGpu.h:252-261).  Hence
//   * an excluded entry whose bitset lies inside another excluded entry's
//     decides no verdict the other does not, in this node or below (both
//     are restricted to the same L' at every push): it is dropped;
//   * a pushable entry whose bitset lies inside an excluded entry's, or whose
//     own group has a member before its pushable one, is non-maximal whenever
//     this node pushes it: with dom_skip it is marked kSkip and not pushed
//     here (the mark is not inherited: below an earlier push it may be
//     maximal).
// The excluded entries tried as dominators are the K of largest popcount
// (ties to the earlier position); an entry equal to a dominator is dropped
// only when it comes later.  The list is compacted in place, in order.
// Returns the new length, the first push (rep << 32 | position) and the
// number of pushable entries not marked.
// One chunk (len <= 64): entries stay in registers, each dominator's bitset
// is read from its lane, and the dominators are chosen and applied round by
// round (a dropped entry is never chosen, and a chosen one is never dropped
// later: a later dominator has no more bits and, if equal, a later lane).
template <int W, bool LDS = false>
__device__ __forceinline__ uint32_t reduce_chunk(const Params& P, Rec<W>* recs, uint32_t len, int K,
                                                 uint32_t minrep, uint64_t& best,
                                                 uint32_t& cbranch, Counters& ctr) {
  const uint32_t lane = lane_id();
  const bool valid = lane < len;
  uint32_t mi = kNone, rp = kNone;
  uint64_t ks = 0, a[W];
#pragma unroll
  for (int w = 0; w < W; ++w) a[w] = 0;
  if (valid) {
    if constexpr (LDS) {
      const auto* q = as_lds(recs) + lane;
      mi = q->minidx & ~kSkip;
      rp = q->rep;
      ks = q->ksum;
#pragma unroll
      for (int w = 0; w < W; ++w) a[w] = q->a[w];
    } else {
      mi = recs[lane].minidx & ~kSkip;
      rp = recs[lane].rep;
      ks = recs[lane].ksum;
#pragma unroll
      for (int w = 0; w < W; ++w) a[w] = recs[lane].a[w];
    }
  }
  const bool excl = valid && rp == kNone && mi < minrep;
  const bool push = valid && rp != kNone;
  uint32_t key = 0;
  if (excl) {
    uint32_t pc = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) pc += __popcll(a[w]);
    key = ((pc + 1) << 6) | (63u - lane);
  }
  // dominators, round by round: the excluded lane of largest key that is
  // neither dropped nor chosen yet
  bool drop = false, chosen = false, sk = P.dom_skip && push && mi < rp;
  for (int r = 0; r < K; ++r) {
    const uint32_t mx = wave_max_u32(excl && !drop && !chosen ? key : 0u);
    if (mx == 0) break;
    const uint32_t pd = 63u - (mx & 63u);
    chosen = chosen || lane == pd;
    bool inside = valid, eq = valid;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const uint64_t dw = read_lane64(a[w], pd);
      inside = inside && (a[w] & ~dw) == 0;
      eq = eq && a[w] == dw;
    }
    // drop excluded lanes inside the dominator but itself and equal ones
    // before it; mark every pushable lane inside it
    drop = drop || (excl && inside && lane != pd && !(eq && lane < pd));
    sk = sk || (push && inside);
  }
  const uint64_t vm = ballot(valid), pm = ballot(push), dm = ballot(drop);
  uint64_t sm = ballot(sk);
  if (!P.dom_skip) sm = 0;
  const uint64_t km = vm & ~dm;
  sm &= km;
  const uint64_t bm = pm & km & ~sm;  // pushes left at this node
  if (dm == 0 && sm == 0) {
    // nothing dropped or marked: the list stands as it is
    if (P.dom_skip) {
      best = bm ? (((uint64_t)read_lane(rp, (uint32_t)__builtin_ctzll(bm)) << 32) |
                   (uint32_t)__builtin_ctzll(bm))
                : ~0ull;
    } else {
      best = wave_min_pair(lane_bit(bm) ? (((uint64_t)rp << 32) | lane) : ~0ull);
    }
    cbranch = __popcll(bm);
    return len;
  }
  const bool keep = lane_bit(km);
  sk = lane_bit(sm);
  const uint32_t pos = lanes_below(km);
  wave_mem_wait();  // every lane has read its entry before any is overwritten
  if (keep) {
    if constexpr (LDS) {
      auto* q = as_lds(recs) + pos;
      q->minidx = mi | (sk ? kSkip : 0u);
      q->rep = rp;
      q->ksum = ks;
#pragma unroll
      for (int w = 0; w < W; ++w) q->a[w] = a[w];
    } else {
      recs[pos].minidx = mi | (sk ? kSkip : 0u);
      recs[pos].rep = rp;
      recs[pos].ksum = ks;
#pragma unroll
      for (int w = 0; w < W; ++w) recs[pos].a[w] = a[w];
    }
  }
  wave_mem_wait();
  if (P.dom_skip) {  // ordered list: the first push is the first such position
    if (bm) {
      const uint32_t l = (uint32_t)__builtin_ctzll(bm);
      best = ((uint64_t)read_lane(rp, l) << 32) | (uint32_t)__popcll(km & ((1ull << l) - 1ull));
    } else {
      best = ~0ull;
    }
  } else {
    best = wave_min_pair(lane_bit(bm) ? (((uint64_t)rp << 32) | pos) : ~0ull);
  }
  cbranch = __popcll(bm);
  if (P.profile) {
    ctr.dropped += __popcll(dm);
    ctr.skipped += __popcll(sm);
  }
  return __popcll(km);
}

template <int W>
__device__ __forceinline__ uint32_t reduce_list(const Params& P, const WaveMem& m, Rec<W>* recs, uint32_t len,
                                uint64_t& best, uint32_t& cbranch, Counters& ctr) {
  const uint32_t lane = lane_id();
  const int K = P.dom_k < kMaxDom ? P.dom_k : kMaxDom;
  if (K <= 0 || best == ~0ull || len < 2 || len < (uint32_t)P.reduce_min || len >= (1u << 20))
    return len;
  const uint64_t tr = P.profile ? cycles() : 0;
  if (len <= (uint32_t)kWave) {
    const uint32_t n =
        in_lds(m, recs)
            ? reduce_chunk<W, true>(P, recs, len, K, (uint32_t)(best >> 32), best, cbranch, ctr)
            : reduce_chunk<W, false>(P, recs, len, K, (uint32_t)(best >> 32), best, cbranch, ctr);
    if (P.profile) ctr.c_reduce += cycles() - tr;
    return n;
  }
  const uint32_t minrep = (uint32_t)(best >> 32);
  lds_flat_fence();  // a list in LDS is reached below through generic pointers
  // dominators: K rounds of a maximum over key = (popcount + 1) << 20 | ~pos
  // among the excluded entries, each round below the previous maximum
  uint32_t dpos[kMaxDom];
  int nd = 0;
  uint32_t prev = 0xFFFFFFFFu;
  for (int r = 0; r < K; ++r) {
    uint32_t kb = 0;
    for (uint32_t base = 0; base < len; base += kWave) {
      const uint32_t i = base + lane;
      if (i < len && recs[i].rep == kNone && (recs[i].minidx & ~kSkip) < minrep) {
        uint32_t pc = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) pc += __popcll(recs[i].a[w]);
        const uint32_t k = ((pc + 1) << 20) | (0xFFFFFu - i);
        if (k < prev && k > kb) kb = k;
      }
    }
    const uint32_t mx = wave_max_u32(kb);
    if (mx == 0) break;
    const uint32_t pd = 0xFFFFFu - (mx & 0xFFFFFu);
    // keep the dominator's bitset: compaction may overwrite its entry
    if (lane < (uint32_t)W) m.dom[nd * kMaxW + lane] = recs[pd].a[lane];
    dpos[nd++] = pd;
    prev = mx;
  }
  lds_flat_fence();
  if (nd == 0 && !P.dom_skip) {
    if (P.profile) ctr.c_reduce += cycles() - tr;
    return len;
  }
  uint32_t out = 0, nbr = 0, ndrop = 0, nskip = 0;
  uint64_t nb = ~0ull;
  for (uint32_t base = 0; base < len; base += kWave) {
    const uint32_t i = base + lane;
    const bool valid = i < len;
    uint32_t mi = kNone, rp = kNone;
    uint64_t ks = 0, a[W];
#pragma unroll
    for (int w = 0; w < W; ++w) a[w] = 0;
    if (valid) {
      mi = recs[i].minidx & ~kSkip;
      rp = recs[i].rep;
      ks = recs[i].ksum;
#pragma unroll
      for (int w = 0; w < W; ++w) a[w] = recs[i].a[w];
    }
    const bool excl = valid && rp == kNone && mi < minrep;
    const bool push = valid && rp != kNone;
    bool drop = false, sk = push && mi < rp;
    for (int d = 0; d < nd; ++d) {
      const uint32_t pd = dpos[d];
      bool inside = true, eq = true;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const uint64_t dw = m.dom[d * kMaxW + w];
        inside = inside && (a[w] & ~dw) == 0;
        eq = eq && a[w] == dw;
      }
      if (inside) {
        drop = drop || (excl && i != pd && !(eq && i < pd));
        sk = sk || push;
      }
    }
    if (!P.dom_skip) sk = false;
    const bool keep = valid && !drop;
    const uint64_t km = ballot(keep);
    const uint32_t pos = out + lanes_below(km);
    wave_mem_wait();  // every lane has read its entry before any is overwritten
    if (keep) {
      recs[pos].minidx = mi | (sk ? kSkip : 0u);
      recs[pos].rep = rp;
      recs[pos].ksum = ks;
#pragma unroll
      for (int w = 0; w < W; ++w) recs[pos].a[w] = a[w];
      if (push && !sk) {
        const uint64_t c = ((uint64_t)rp << 32) | pos;
        nb = c < nb ? c : nb;
      }
    }
    nbr += __popcll(ballot(keep && push && !sk));
    nskip += __popcll(ballot(keep && sk));
    ndrop += __popcll(ballot(valid && drop));
    out += __popcll(km);
  }
  lds_flat_fence();
  best = (wave_min_pair(nb));
  cbranch = nbr;
  if (P.profile) {
    ctr.dropped += ndrop;
    ctr.skipped += nskip;
    ctr.c_reduce += cycles() - tr;
  }
  return out;
}

// Whether sharing a node is worthwhile now: the top-level vertices are all
// claimed, the queue is short and one of this wave's context slots is free
// and large enough.  Returns that slot's id, or kNone.
__device__ __forceinline__ uint32_t share_wanted(const Params& P, uint64_t gw, uint32_t nL,
                                                 uint32_t nC, uint32_t len, int W) {
  const uint64_t bytes = sizeof(CtxHdr) + (uint64_t)len * (16 + 8 * W) + (uint64_t)nL * 8 +
                         (P.count_bytes ? (uint64_t)nC * 4 : 0);
  if (bytes > P.w.ctx_bytes) return kNone;
  if (uni64(ld_dev(P.s.root_next)) < roots_total(P.s)) return kNone;
  const long long depth = (long long)uni64(ld_dev(P.s.queued));
  if (depth >= (long long)P.share_queue) return kNone;
  return free_ctx(P, gw);
}

template <int W, bool KL_R, bool C_R>
__device__ __forceinline__ void dfs(const Params& P, const WaveMem& m, uint32_t nL, uint32_t nC, Frame F,
                    uint32_t& lds_top, uint64_t& hbm_top, Counters& ctr, const TaskArrays& ta,
                    uint64_t root_rec, uint32_t root_len, uint32_t nR0, uint64_t gw, uint32_t fb);

// What to do with a child node (list at crec, committed in the arenas): run it
// with re-indexed one-word bitsets when its L fits one word and the task's
// does not, merge twins, share it with other waves when it is large and work
// is scarce.  Returns true when the caller should descend into (crec, clen)
// itself; false when the node has been handled here.
template <int W, bool KL_R, bool C_R>
__device__ __forceinline__ bool child_node(const Params& P, const WaveMem& m, uint64_t gw, Counters& ctr,
                           uint32_t nL, uint32_t nC, const TaskArrays& ta, Rec<W>* crec,
                           uint32_t& clen, uint64_t& best_child, uint32_t& cbranch, uint32_t nb,
                           const uint64_t* ax, uint64_t sumR, uint32_t& lds_top,
                           uint64_t& hbm_top, uint32_t fb_next) {
  const uint32_t lane = lane_id();
  if constexpr (W > 1) {
    if (nb <= 64 && !P.list) {
      const uint32_t sl = lds_top;
      const uint64_t sh = hbm_top;
      uint32_t kt = 1, dt = 1;
      uint64_t ktop = 0, dtop = 0;
      uint64_t kaddr = arena_take(m, lds_top, hbm_top, 64 * 8, kt, ktop);
      if (kaddr) {
        if (kt == 0) lds_top = (uint32_t)(ktop + 512); else hbm_top = ktop + 512;
      }
      uint64_t daddr = kaddr ? arena_take(m, lds_top, hbm_top, (uint64_t)clen * sizeof(Rec<1>),
                                          dt, dtop) : 0;
      if (daddr) {
        if (dt == 0) lds_top = (uint32_t)(dtop + (uint64_t)clen * sizeof(Rec<1>));
        else hbm_top = dtop + (uint64_t)clen * sizeof(Rec<1>);
        uint64_t* kLc = (uint64_t*)kaddr;
        Rec<1>* dst = (Rec<1>*)daddr;
        lds_flat_fence();  // ta.kL and crec are read below through generic pointers
        // keys of the child's L in bit order
        {
          uint32_t t = 0;
          for (int w = 0; w < W; ++w) {
            const uint64_t mw = ax[w];
            const bool b = (mw >> lane) & 1ull;
            const uint32_t pos = t + __popcll(mw & ((1ull << lane) - 1ull));
            if (b) kLc[pos] = rd<KL_R>(ta.kL + w * 64 + lane);
            t += __popcll(mw);
          }
        }
        uint64_t tc0 = P.profile ? cycles() : 0;
        compress_list<W>(crec, clen, ax, dst);
        if (P.profile) ctr.c_compress += cycles() - tc0;
        uint32_t dlen = clen, dcb = cbranch;
        uint64_t dbest = best_child;
        if (P.group && dlen >= (uint32_t)P.merge_min && dcb >= (uint32_t)P.merge_branch) {
          uint64_t tm = P.profile ? cycles() : 0;
          dlen = merge_twins<1>(m, dst, dlen, lds_top, hbm_top, dbest, dcb, P.dom_skip != 0);
          if (P.profile) ctr.c_merge += cycles() - tm;
          ctr.merges++;
        }
        dlen = reduce_list<1>(P, m, dst, dlen, dbest, dcb, ctr);
        if (dcb == 0) {  // every push of the node is known to be non-maximal
          lds_top = sl;
          hbm_top = sh;
          return false;
        }
        TaskArrays tc{kLc, ta.C};
        bool shared = false;
        const uint32_t cx = P.share && (uint64_t)dlen * dcb >= (uint64_t)P.share_work
                                ? share_wanted(P, gw, nb, nC, dlen, 1)
                                : kNone;
        if (cx != kNone) {
          uint64_t tq = P.profile ? cycles() : 0;
          shared = publish_list<1, false, C_R>(P, gw, cx, graph_id(P, m), dst, dlen, nb, nC, nb, sumR, tc);
          if (P.profile) ctr.c_publish += cycles() - tq;
        }
        if (shared) {
          ctr.splits++;
          ctr.publishes++;
        } else {
          if (dt == 0) lds_top = (uint32_t)(dtop + (uint64_t)dlen * sizeof(Rec<1>));
          else hbm_top = dtop + (uint64_t)dlen * sizeof(Rec<1>);
          Frame G;
          G.sumR = sumR;
          G.rec = daddr;
          G.len = dlen;
          G.cur = (uint32_t)dbest;
          G.curidx = (uint32_t)(dbest >> 32);
          G.nbits = nb;
          G.tier = dt;
          G.top = (uint32_t)dtop;
          dfs<1, false, C_R>(P, m, nb, nC, G, lds_top, hbm_top, ctr, tc, 0, 0, 0, gw, fb_next);
        }
        lds_top = sl;
        hbm_top = sh;
        return false;
      }
      lds_top = sl;
      hbm_top = sh;
    }
  }
  if (P.group && !P.list && clen >= (uint32_t)P.merge_min &&
      cbranch >= (uint32_t)P.merge_branch) {
    uint64_t tm = P.profile ? cycles() : 0;
    clen = merge_twins<W>(m, crec, clen, lds_top, hbm_top, best_child, cbranch, P.dom_skip != 0);
    if (P.profile) ctr.c_merge += cycles() - tm;
    ctr.merges++;
  }
  clen = reduce_list<W>(P, m, crec, clen, best_child, cbranch, ctr);
  if (cbranch == 0) return false;  // every push of the node is known to be non-maximal
  const uint32_t cx = P.share && !P.list && (uint64_t)clen * cbranch >= (uint64_t)P.share_work
                          ? share_wanted(P, gw, nL, nC, clen, W)
                          : kNone;
  if (cx != kNone) {
    uint64_t tq = P.profile ? cycles() : 0;
    bool ok = publish_list<W, KL_R, C_R>(P, gw, cx, graph_id(P, m), crec, clen, nL, nC, nb, sumR, ta);
    if (P.profile) ctr.c_publish += cycles() - tq;
    if (ok) {
      ctr.splits++;
      ctr.publishes++;
      return false;
    }
  }
  return true;
}

// Late sharing: when other waves run dry while this one is deep in a large
// subtree, hand the unpushed rest of its shallowest open frame to them.  The
// frame's remaining pushes are exactly its entries with candidate >= its next
// candidate (pushes go by increasing candidate and pruned entries are marked
// unpushable in place), so publishing that suffix and closing the frame keeps
// the traversal's result.  Returns true when the current frame was closed.
template <int W, bool KL_R, bool C_R>
__device__ __forceinline__ bool late_share(const Params& P, const WaveMem& m, uint64_t gw, uint32_t nL,
                           uint32_t nC,
                           const TaskArrays& ta, Frame& F, uint32_t fb, uint32_t depth,
                           Counters& ctr) {
  // cheap gate first: the top-level vertices are all claimed and the queue is
  // short; then pick the shallowest frame with pushes left
  if (uni64(ld_dev(P.s.root_next)) < roots_total(P.s)) return false;
  if ((long long)uni64(ld_dev(P.s.queued)) >= (long long)P.share_queue) return false;
  uint32_t at = kNone;
  Frame S;
  const uint32_t scan = depth < 4u ? depth : 4u;
  for (uint32_t i = 0; i < scan; ++i) {
    const Frame Fi = frame_get(m, fb + i);
    if (Fi.cur != kNone) {
      at = i;
      S = Fi;
      break;
    }
  }
  if (at == kNone) {
    if (depth > 4u) return false;
    S = F;
  }
  if (S.len < (uint32_t)P.late_share) return false;
  const uint32_t cx = share_wanted(P, gw, nL, nC, S.len, W);
  if (cx == kNone) return false;
  uint64_t tq = P.profile ? cycles() : 0;
  const bool ok = publish_list<W, KL_R, C_R>(P, gw, cx, graph_id(P, m), (const Rec<W>*)S.rec, S.len, nL, nC,
                                             S.nbits, S.sumR, ta, S.curidx, 2);
  if (P.profile) ctr.c_publish += cycles() - tq;
  if (!ok) return false;
  ctr.splits++;
  ctr.publishes++;
  ctr.late++;
  if (at == kNone) {
    F.cur = kNone;
    return true;
  }
  S.cur = kNone;
  frame_put(m, fb + at, S);
  return false;
}

// The stack machine from frame F (a list in this wave's arenas); saved
// frames go to fr[0..).
template <int W, bool KL_R, bool C_R>
__device__ __forceinline__ void dfs(const Params& P, const WaveMem& m, uint32_t nL, uint32_t nC, Frame F,
                    uint32_t& lds_top, uint64_t& hbm_top, Counters& ctr, const TaskArrays& ta,
                    uint64_t root_rec, uint32_t root_len, uint32_t nR0, uint64_t gw, uint32_t fb) {
  typedef Rec<W> R_t;
  const uint32_t lane = lane_id();
  const bool prune = P.prune != 0;
  const int64_t* voff = m.g->v_off;
  const uint32_t swapped = m.g->swapped;
  uint32_t depth = 0, tick = 0;
  // this subtree's result counters, in scalar registers until it ends
  uint64_t l_nodes = 0, l_nonmax = 0, l_max = 0, l_hsum = 0, l_hxor = 0;
  auto flush = [&]() {
    ctr.nodes += l_nodes;
    ctr.nonmax += l_nonmax;
    ctr.maximal += l_max;
    ctr.hsum += l_hsum;
    ctr.hxor ^= l_hxor;
  };
  for (;;) {
    ++tick;
    // one wall-clock read per push: a push over a long list costs far more
    if (P.budget != 0 && past_deadline(P, m.deadline)) {
      ctr.abandoned++;
      graph_work(P, graph_id(P, m), 1);  // never finished: the graph stays incomplete
      flush();
      return;
    }
    if (F.cur == kNone) {
      // frame exhausted: free its list, return to the parent
      if (F.tier == 0) lds_top = F.top; else hbm_top = F.top;
      if (depth == 0) break;
      --depth;
      F = frame_get(m, fb + depth);
      continue;
    }
    if (P.late_share && !P.list && (tick & 31u) == 0 &&
        late_share<W, KL_R, C_R>(P, m, gw, nL, nC, ta, F, fb, depth, ctr))
      continue;  // the current frame was handed over: F.cur is now kNone
    R_t* recs = (R_t*)F.rec;
    const uint32_t xpos = F.cur, xidx = F.curidx;
    if (P.count_bytes) {
      // B(n) = 4 (deg(x) + |L_parent| + |C|)
      uint32_t xc = rd<C_R>(ta.C + xidx);
      uint64_t dx = (uint64_t)(voff[xc + 1] - voff[xc]);
      ctr.balg += 4ull * (dx + F.nbits + nC);
    }
    uint64_t ax[W];
    uint32_t nb = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      ax[w] = uni64(F.tier == 0 ? as_lds(recs)[xpos].a[w] : recs[xpos].a[w]);
      nb += __popcll(ax[w]);
    }
    // speculative child list (at most as long as the parent's)
    uint32_t ctier = 1;
    uint64_t ctop = 0;
    uint64_t crec_addr =
        arena_take(m, lds_top, hbm_top, (uint64_t)F.len * sizeof(R_t), ctier, ctop);
    if (crec_addr == 0) {
      // no room for the child list: report and abandon the task
      ctr.overflow++;
      ctr.ovf_arena++;
      flush();
      return;
    }
    R_t* crec = (R_t*)crec_addr;
    uint64_t tp = P.profile ? cycles() : 0;
    const bool ord = P.dom_skip != 0;
    PassOut o;
    if constexpr (W == 1) {
      if (F.tier == 0 && ctier == 0)
        o = push_pass<1, false, true, true>(recs, F.len, ax, xidx, crec, prune, ord);
      else if (F.tier == 0)
        o = push_pass<1, false, true, false>(recs, F.len, ax, xidx, crec, prune, ord);
      else if (ctier == 0)
        o = push_pass<1, false, false, true>(recs, F.len, ax, xidx, crec, prune, ord);
      else
        o = push_pass<1, false, false, false>(recs, F.len, ax, xidx, crec, prune, ord);
    } else {
      // generic (flat_) accesses: fenced against the ds_ accesses of LDS lists
      const bool lds_any = F.tier == 0 || ctier == 0;
      if (lds_any) lds_flat_fence();
      o = push_pass<W, false>(recs, F.len, ax, xidx, crec, prune, ord);
      if (lds_any) lds_flat_fence();
    }
    if (P.profile) ctr.c_pass += cycles() - tp;
    ctr.nodes++;
    if (P.count_bytes) ctr.blist += (uint64_t)(F.len + o.clen) * sizeof(R_t);
    if (P.profile) {
      ctr.entries += F.len;
      if (W == 1) {
        ctr.ent_w1 += F.len;
        ctr.nodes_w1++;
      }
    }
    F.cur = o.best_next == ~0ull ? kNone : (uint32_t)o.best_next;
    F.curidx = (uint32_t)(o.best_next >> 32);
    if (o.nonmax) {
      ctr.nonmax++;
      continue;
    }
    const uint64_t sumL = key_sum_of_bits<W, KL_R>(m, ax, ta);
    const uint64_t sumR = F.sumR + o.absorbed;
    add_biclique(ctr, biclique_hash(sumL, sumR, swapped));
    if constexpr (!KL_R && !C_R) {
      if (P.list) {
        // R' = R0 ∪ {C[i] : a_i ⊇ L'} over the (ungrouped) root list
        const R_t* root = (const R_t*)root_rec;
        uint32_t nr = nR0;
        for (uint32_t base = 0; base < root_len; base += kWave) {
          uint32_t i = base + lane;
          bool s = false;
          if (i < root_len) {
            s = true;
#pragma unroll
            for (int w = 0; w < W; ++w) s = s && ((root[i].a[w] & ax[w]) == ax[w]);
          }
          nr += __popcll(ballot(s));
        }
        uint64_t at = list_reserve(P, 2ull + nb + nr);
        if (at != ~0ull) {
          uint32_t* out = P.list + at;
          if (lane == 0) {
            out[0] = nb;
            out[1] = nr;
          }
          // L' ids: the task L is kept in m.T in list mode
          uint32_t c = 0;
          for (int w = 0; w < W; ++w) {
            bool b = (ax[w] >> lane) & 1ull;
            uint64_t bmk = ballot(b);
            if (b) out[2 + c + lanes_below(bmk)] = m.T[w * 64 + lane];
            c += __popcll(bmk);
          }
          for (uint32_t i = lane; i < nR0; i += kWave) out[2 + nb + i] = m.R[i];
          c = nR0;
          for (uint32_t base = 0; base < root_len; base += kWave) {
            uint32_t i = base + lane;
            bool s = false;
            if (i < root_len) {
              s = true;
#pragma unroll
              for (int w = 0; w < W; ++w) s = s && ((root[i].a[w] & ax[w]) == ax[w]);
            }
            uint64_t sm = ballot(s);
            if (s) out[2 + nb + c + lanes_below(sm)] = m.C[root[i].minidx & ~kSkip];
            c += __popcll(sm);
          }
        }
      }
    }
    if (nb > 1 && o.best_child != ~0ull) {
      // commit the child list, then decide how to run it
      const uint32_t sl = lds_top;
      const uint64_t sh = hbm_top;
      if (ctier == 0) lds_top = (uint32_t)(ctop + (uint64_t)o.clen * sizeof(R_t));
      else hbm_top = ctop + (uint64_t)o.clen * sizeof(R_t);
      bool here = child_node<W, KL_R, C_R>(P, m, gw, ctr, nL, nC, ta, crec, o.clen, o.best_child,
                                           o.cbranch, nb, ax, sumR, lds_top, hbm_top,
                                           fb + depth);
      lds_top = sl;
      hbm_top = sh;
      if (!here) continue;
      // descend: the child list becomes a frame
      frame_put(m, fb + depth, F);
      ++depth;
      if (ctier == 0) lds_top = (uint32_t)(ctop + (uint64_t)o.clen * sizeof(R_t));
      else hbm_top = ctop + (uint64_t)o.clen * sizeof(R_t);
      F.sumR = sumR;
      F.rec = crec_addr;
      F.len = o.clen;
      F.cur = (uint32_t)o.best_child;
      F.curidx = (uint32_t)(o.best_child >> 32);
      F.nbits = nb;
      F.tier = ctier;
      F.top = (uint32_t)ctop;
    }
  }
  flush();
}

// Run the items of one entry (`word`: lane l holds word l).
template <int W>
__device__ __forceinline__ void run_items_w(const Params& P, const WaveMem& m, CtxHdr* h, uint32_t nL,
                            uint32_t nC, uint32_t len, uint32_t nbits, uint64_t sumR0,
                            uint32_t word, uint32_t cnt, Counters& ctr, uint64_t gw) {
  typedef Rec<W> R_t;
  R_t* recs;
  TaskArrays rta;
  ctx_arrays<W>(h, nL, nC, len, recs, rta);
  // the hash keys of L, read at every maximal push below: a local copy at
  // the bottom of the arenas (LDS when it fits)
  uint32_t lds0 = 0;
  uint64_t hbm0 = 0;
  uint32_t kt = 1;
  uint64_t ktop = 0;
  const uint64_t kbytes = ((uint64_t)nL * 8 + 15) & ~15ull;
  uint64_t kaddr = arena_take(m, lds0, hbm0, kbytes, kt, ktop);
  if (kaddr == 0) {
    ctr.overflow += cnt;
    ctr.ovf_arena += cnt;
    if (lane_id() == 0)
      __hip_atomic_fetch_add(&h->refs, (unsigned long long)(-(long long)cnt), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  if (kt == 0) lds0 = (uint32_t)kbytes; else hbm0 = kbytes;
  for (uint32_t j = lane_id(); j < nL; j += kWave) {
    const uint64_t key = ld_dev(rta.kL + j);
    if (kt == 0) as_lds((uint64_t*)kaddr)[j] = key;
    else ((uint64_t*)kaddr)[j] = key;
  }
  wave_mem_wait();
  const TaskArrays ta{(const uint64_t*)kaddr, rta.C};
  for (uint32_t it = 0; it < cnt; ++it) {
    const uint32_t xpos = read_lane(word, 2 + 2 * it), xidx = read_lane(word, 3 + 2 * it);
    uint64_t t0 = P.profile ? cycles() : 0;
    if (P.count_bytes) {
      uint32_t xc = ld_dev(ta.C + xidx);
      uint64_t dx = (uint64_t)(m.g->v_off[xc + 1] - m.g->v_off[xc]);
      ctr.balg += 4ull * (dx + nbits + nC);
    }
    uint64_t ax[W];
    uint32_t nb = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      ax[w] = uni64(ld_dev(&recs[xpos].a[w]));
      nb += __popcll(ax[w]);
    }
    uint32_t lds_top = lds0;
    uint64_t hbm_top = hbm0;
    uint32_t ctier = 1;
    uint64_t ctop = 0;
    uint64_t crec_addr = arena_take(m, lds_top, hbm_top, (uint64_t)len * sizeof(R_t), ctier, ctop);
    if (crec_addr == 0) {
      ctr.overflow++;
      ctr.ovf_arena++;
      continue;
    }
    uint64_t tp = P.profile ? cycles() : 0;
    PassOut o = push_pass<W, true>(recs, len, ax, xidx, (R_t*)crec_addr, false, P.dom_skip != 0);
    if (ctier == 0) lds_flat_fence();  // the child list went to LDS through flat_ stores
    if (P.profile) ctr.c_pass += cycles() - tp;
    ctr.nodes++;
    if (P.count_bytes) ctr.blist += (uint64_t)(len + o.clen) * sizeof(R_t);
    ctr.entries += len;
    if (W == 1) {
      ctr.ent_w1 += len;
      ctr.nodes_w1++;
    }
    if (o.nonmax) {
      ctr.nonmax++;
      phase(ctr, P.profile, 5, t0);
      continue;
    }
    const uint64_t sumL = key_sum_of_bits<W, false>(m, ax, ta);
    const uint64_t sumR = sumR0 + o.absorbed;
    add_biclique(ctr, biclique_hash(sumL, sumR, m.g->swapped));
    if (nb > 1 && o.best_child != ~0ull) {
      if (ctier == 0) lds_top = (uint32_t)(ctop + (uint64_t)o.clen * sizeof(R_t));
      else hbm_top = ctop + (uint64_t)o.clen * sizeof(R_t);
      bool here = child_node<W, false, true>(P, m, gw, ctr, nL, nC, ta, (R_t*)crec_addr, o.clen,
                                             o.best_child, o.cbranch, nb, ax, sumR, lds_top,
                                             hbm_top, 0);
      if (here) {
        if (ctier == 0) lds_top = (uint32_t)(ctop + (uint64_t)o.clen * sizeof(R_t));
        else hbm_top = ctop + (uint64_t)o.clen * sizeof(R_t);
        Frame F;
        F.sumR = sumR;
        F.rec = crec_addr;
        F.len = o.clen;
        F.cur = (uint32_t)o.best_child;
        F.curidx = (uint32_t)(o.best_child >> 32);
        F.nbits = nb;
        F.tier = ctier;
        F.top = (uint32_t)ctop;
        dfs<W, false, true>(P, m, nL, nC, F, lds_top, hbm_top, ctr, ta, 0, 0, 0, gw, 0);
      }
    }
    phase(ctr, P.profile, 5, t0);
  }
  if (lane_id() == 0)
    __hip_atomic_fetch_add(&h->refs, (unsigned long long)(-(long long)cnt), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

template <int WMAX>
__device__ __forceinline__ void run_items(const Params& P, const WaveMem& m, uint32_t word, Counters& ctr,
                          uint64_t gw) {
  const uint32_t cnt = read_lane(word, 0) & ~kItemKind;
  const uint32_t owner = read_lane(word, 1);
  CtxHdr* h = ctx_hdr(P, owner);
  const uint32_t nL = uni(ld_dev(&h->nL)), nC = uni(ld_dev(&h->nC)), len = uni(ld_dev(&h->len)),
                 W = uni(ld_dev(&h->W));
  const uint32_t nbits = uni(ld_dev(&h->nbits));
  const uint64_t sumR0 = uni64(ld_dev(&h->sumR0));
  if (W == 1) run_items_w<1>(P, m, h, nL, nC, len, nbits, sumR0, word, cnt, ctr, gw);
  else if (W == 2) run_items_w<2>(P, m, h, nL, nC, len, nbits, sumR0, word, cnt, ctr, gw);
  else if (W == 4) run_items_w<4>(P, m, h, nL, nC, len, nbits, sumR0, word, cnt, ctr, gw);
  else if constexpr (WMAX >= 16) {
    if (W == 8) run_items_w<8>(P, m, h, nL, nC, len, nbits, sumR0, word, cnt, ctr, gw);
    else run_items_w<16>(P, m, h, nL, nC, len, nbits, sumR0, word, cnt, ctr, gw);
  }
}

// Build the root list of a task (entry i starts as candidate i alone:
// minidx = i, rep = i if i >= next_id else kNone, bitset of L ∩ N(C[i]) with
// Wt words, zero-padded beyond |L|; with one-word bitsets the candidates are
// then grouped by bitset through a per-wave hash table), then either publish
// it as a shared context (share) or run the stack machine here.
template <int Wt>
__device__ __forceinline__ void enumerate_task(const Params& P, const WaveMem& m, const uint32_t* L, uint32_t nL,
                               uint32_t nC, uint32_t next_id, uint64_t sumR, uint32_t nR,
                               Counters& ctr, uint64_t& t0, uint32_t share_ctx, uint64_t gw) {
  typedef Rec<Wt> R_t;
  const DevGraph g = *m.g;
  const uint32_t lane = lane_id();
  int32_t* cmap = m.cmap;
  uint32_t lds_top = 0;
  uint64_t hbm_top = 0;
  // candidate bitsets, one record per candidate, built in the HBM arena
  if ((uint64_t)nC * sizeof(R_t) > m.arena_bytes) {
    ctr.overflow++;
    ctr.ovf_arena++;
    for (uint32_t i = lane; i < nC; i += kWave) st_wg(cmap + m.C[i], 0);
    wave_mem_wait();
    return;
  }
  const uint64_t braw = (uint64_t)m.arena;
  hbm_top = (uint64_t)nC * sizeof(R_t);
  if (P.count_bytes) ctr.blist += (uint64_t)nC * sizeof(R_t);
  R_t* brec = (R_t*)braw;
  for (uint32_t i = lane; i < nC; i += kWave) {
    brec[i].minidx = i;
    brec[i].rep = i >= next_id ? i : kNone;
    brec[i].ksum = m.kC[i];
#pragma unroll
    for (int w = 0; w < Wt; ++w) brec[i].a[w] = 0ull;
  }
  wave_mem_wait();
  for_each_edge(L, nL, g.u_off, g.u_adj, m.s_off, m.s_start, [&](uint32_t j, uint32_t w) {
    int32_t p = ld_wg(cmap + w);
    if (p < 0) or_wg(&brec[-p - 1].a[j >> 6], 1ull << (j & 63));
  });
  wave_mem_wait();
  for (uint32_t i = lane; i < nC; i += kWave) st_wg(cmap + m.C[i], 0);
  wave_mem_wait();
  // group twins (not in list mode: its R' lists need one entry per candidate)
  uint32_t len = nC;
  const uint64_t list_addr = braw;
  uint64_t best = ~0ull;
  uint32_t nbr = 0;
  if (P.group && !P.list && nC > 1) {
    len = merge_twins<Wt>(m, brec, nC, lds_top, hbm_top, best, nbr, P.dom_skip != 0);
  }
  if (len == nC) {
    // first push: the smallest rep
    for (uint32_t i = lane; i < len; i += kWave) {
      const uint32_t rp = brec[i].rep;
      if (rp != kNone) {
        const uint64_t c = ((uint64_t)rp << 32) | i;
        best = c < best ? c : best;
      }
    }
    best = (wave_min_pair(best));
  }
  // drop dominated excluded candidates and mark the non-maximal pushes (not in
  // list mode: its R' lists read the root list)
  if (!P.list) len = reduce_list<Wt>(P, m, brec, len, best, nbr, ctr);
  wave_mem_wait();
  if (best == ~0ull) {  // nothing to push below the task's own biclique
    phase(ctr, P.profile, 5, t0);
    return;
  }
  TaskArrays lta{m.kL, m.C};
  const uint64_t need = sizeof(CtxHdr) + (uint64_t)len * sizeof(R_t) + (uint64_t)nL * 8 +
                        (P.count_bytes ? (uint64_t)nC * 4 : 0);
  if (share_ctx != kNone && need <= P.w.ctx_bytes &&
      publish_list<Wt, false, false>(P, gw, share_ctx, graph_id(P, m), (const R_t*)list_addr, len, nL, nC, nL,
                                     sumR, lta)) {
    ctr.splits++;
    phase(ctr, P.profile, 6, t0);
    return;
  }
  // the hash keys of L, read at every maximal push: into LDS when small
  const uint64_t* kL = m.kL;
  uint32_t lds0 = 0;
  const uint32_t kbytes = (nL * 8 + 15) & ~15u;
  if (kbytes <= m.lds_bytes / 4) {
    for (uint32_t j = lane; j < nL; j += kWave) as_lds((uint64_t*)m.lds)[j] = m.kL[j];
    kL = (const uint64_t*)m.lds;
    lds0 = kbytes;
  }
  // final root list: into LDS when it fits, else at the HBM arena's base
  uint64_t root = list_addr;
  uint32_t rtier = 1;
  if ((uint64_t)lds0 + (uint64_t)len * sizeof(R_t) <= m.lds_bytes) {
    R_t* dst = (R_t*)(m.lds + lds0);
    const R_t* src = (const R_t*)list_addr;
    for (uint32_t i = lane; i < len; i += kWave) {
      auto* q = as_lds(dst) + i;
      q->minidx = src[i].minidx;
      q->rep = src[i].rep;
      q->ksum = src[i].ksum;
#pragma unroll
      for (int w = 0; w < Wt; ++w) q->a[w] = src[i].a[w];
    }
    root = (uint64_t)dst;
    rtier = 0;
    lds_top = lds0 + (uint32_t)((uint64_t)len * sizeof(R_t));
    hbm_top = 0;
  } else {
    lds_top = lds0;
    hbm_top = (uint64_t)len * sizeof(R_t);
  }
  wave_mem_wait();
  phase(ctr, P.profile, 4, t0);
  {
    Frame F;
    F.sumR = sumR;
    F.rec = root;
    F.len = len;
    F.cur = (uint32_t)best;
    F.curidx = (uint32_t)(best >> 32);
    F.nbits = nL;
    F.tier = rtier;
    F.top = rtier == 0 ? lds0 : 0;
    TaskArrays ta{kL, m.C};
    dfs<Wt, false, false>(P, m, nL, nC, F, lds_top, hbm_top, ctr, ta, root, len, nR, gw, 0);
  }
  phase(ctr, P.profile, 5, t0);
}

// ---------------------------------------------------------------------------
// one task

// Common neighbourhood of a prefix S[0..kp) (the L of a split parent), shared
// by all children of one queue entry.  kp = 1 points into the CSR row.
struct Prefix {
  const uint32_t* L;
  uint32_t n;
  uint64_t sdeg;  // sum of the prefix vertices' degrees
};

__device__ __forceinline__ Prefix prefix_common(const DevGraph& g, const WaveMem& m, const Task& t, int kp) {
  const uint32_t lane = lane_id();
  Prefix p;
  uint32_t v0 = t.v[0];
  p.sdeg = (uint64_t)(g.v_off[v0 + 1] - g.v_off[v0]);
  p.n = (uint32_t)(g.v_off[v0 + 1] - g.v_off[v0]);
  const uint32_t* cur = g.v_adj + g.v_off[v0];
  for (int i = 1; i < kp; ++i) {
    uint32_t vi = t.v[i];
    uint32_t di = (uint32_t)(g.v_off[vi + 1] - g.v_off[vi]);
    p.sdeg += di;
    uint32_t* dst = (i & 1) ? m.T : m.Lp;
    p.n = wave_intersect(cur, p.n, g.v_adj + g.v_off[vi], di, dst);
    cur = dst;
  }
  if (kp > 1 && cur != m.Lp) {
    for (uint32_t i = lane; i < p.n; i += kWave) m.Lp[i] = cur[i];
    wave_sync();
    cur = m.Lp;
  }
  p.L = cur;
  return p;
}

// One task S = t.v[0..k).  For k > 1, `pre` is the common neighbourhood of
// S[0..k-1) (prefix_common).
template <int WMAX>
__device__ __forceinline__ void run_task(const Params& P, const WaveMem& m, const Task& t, int k,
                         const Prefix& pre, Counters& ctr, uint64_t gw) {
  const DevGraph g = *m.g;
  const uint32_t lane = lane_id();
  ctr.tasks++;
  ctr.nodes++;
  uint64_t t0 = P.profile ? cycles() : 0;
  // L = ∩ N(S) (get_common_neighbors, src/IterFinderGpu.cu:680-711)
  const uint32_t* L;
  uint32_t nL;
  const uint32_t np = k > 1 ? pre.n : 0;
  const uint32_t vk = t.v[k - 1];
  const uint32_t dk = (uint32_t)(g.v_off[vk + 1] - g.v_off[vk]);
  uint64_t sdeg = dk;  // sum of the task vertices' degrees
  if (k == 1) {
    L = g.v_adj + g.v_off[vk];
    nL = dk;
  } else {
    sdeg += pre.sdeg;
    nL = wave_intersect(pre.L, pre.n, g.v_adj + g.v_off[vk], dk, m.L);
    L = m.L;
  }
  if (nL == 0) {  // isolated vertex: no biclique (undefined in the reference, Q5)
    if (P.count_bytes) {
      ctr.balg += 4ull * sdeg;
      ctr.blist += 4ull * sdeg;
    }
    ctr.nonmax++;
    return;
  }
  phase(ctr, P.profile, 1, t0);
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
  if (P.count_bytes) {
    ctr.balg += 4ull * (sdeg + root_edges);
    ctr.blist += 4ull * (sdeg + root_edges);
  }
  ctr.edges += root_edges;
  wave_mem_wait();
  phase(ctr, P.profile, 2, t0);
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
    // every w in R below vk outside the prefix must contain the prefix's L
    canonical = true;
    for (uint32_t base = 0; base < nR && canonical; base += kWave) {
      uint32_t i = base + lane;
      uint32_t w = i < nR ? m.R[i] : kNone;
      bool below = w < vk;
      bool in_prefix = false;
      for (int j = 0; j < k - 1; ++j) in_prefix |= (t.v[j] == w);
      uint64_t cand = ballot(below && !in_prefix);
      while (cand && canonical) {
        uint32_t wc = read_lane(w, first_lane(cand));
        cand &= cand - 1;
        uint32_t c = wave_intersect_count(pre.L, np, g.v_adj + g.v_off[wc],
                                          (uint32_t)(g.v_off[wc + 1] - g.v_off[wc]));
        canonical = uni(c) == np;
      }
      if (!ballot(below)) break;
      if (ballot(!below)) break;
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
  phase(ctr, P.profile, 3, t0);
  if (nL <= 1 || rest == 0) {
    count_root(sumL, sumR);
    clear_cmap();
    return;
  }
  const uint32_t W = (nL + 63) / 64;
  // split (src/IterFinderGpu.cu:835-858)
  bool rule = P.split && ((rest > (uint32_t)g.bound_height || nL > (uint32_t)g.bound_height) &&
                          (uint64_t)rest * nL > (uint64_t)g.bound_size);
  bool must = W > (uint32_t)kMaxW;  // |L| beyond any bitset: split regardless
  // large task: share its root list with other waves when one of this wave's
  // context slots is free and large enough
  uint32_t share_ctx = kNone;
  // (whether its reduced root list fits the slot is decided once it is built)
  if (rule && P.share && !P.list && W <= (uint32_t)WMAX) share_ctx = free_ctx(P, gw);
  const bool share = share_ctx != kNone;
  // a large task with no context slot free: its children are recomputed by
  // other waves (recompute, the reference's LargeToTiny) or, by default, the
  // task runs here and late sharing hands its frames out when waves run dry
  const bool recompute = rule && k < P.max_prefix && (P.recompute || !P.share || P.list);
  if (!share && (recompute || (must && k < kMaxPrefix))) {
    // children S + [C[i]], i >= next_id, each recomputed by its wave
    const unsigned long long ne = (rest + kChunk - 1) / kChunk;
    if (lane == 0)
      __hip_atomic_fetch_add(P.s.work_left, ne, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    graph_work(P, t.gid, (long long)ne);
    uint32_t pushed = 0;
    if (queue_push_children(P.s, t, k, m.C + next_id, rest, pushed,
                            (uint32_t)(gw & (kShards - 1)))) {
      ctr.splits++;
      count_root(sumL, sumR);
      clear_cmap();
      phase(ctr, P.profile, 6, t0);
      return;
    }
    if (lane == 0)
      __hip_atomic_fetch_add(P.s.work_left, (unsigned long long)(-(long long)ne),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    graph_work(P, t.gid, -(long long)ne);
  }
  if (must) {
    ctr.overflow++;
    ctr.ovf_wide++;
    clear_cmap();
    return;
  }
  if (W > (uint32_t)WMAX) {
    // too wide for this kernel: hand the whole task to the wide-bitset pass
    clear_cmap();
    ctr.tasks--;
    ctr.nodes--;
    if (P.count_bytes) ctr.balg -= 4ull * (sdeg + root_edges);
    unsigned long long at = ~0ull;
    if (lane == 0) {
      at = __hip_atomic_fetch_add(P.s.n_deferred, 1ull, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
      if (at >= P.s.deferred_cap) at = ~0ull;
    }
    at = uni64(at);
    if (at == ~0ull) {
      ctr.overflow++;
      ctr.ovf_defer++;
      return;
    }
    if (lane < (uint32_t)kMaxPrefix) st_dev(&P.s.deferred[at].v[lane], t.v[lane]);
    if (lane == (uint32_t)kMaxPrefix) st_dev(&P.s.deferred[at].gid, t.gid);
    graph_work(P, t.gid, 1);  // the wide pass finishes it
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
    enumerate_task<1>(P, m, L, nL, nC, next_id, sumR, nR, ctr, t0, share_ctx, gw);
  } else if (W <= 2) {
    enumerate_task<2>(P, m, L, nL, nC, next_id, sumR, nR, ctr, t0, share_ctx, gw);
  } else if (W <= 4) {
    enumerate_task<4>(P, m, L, nL, nC, next_id, sumR, nR, ctr, t0, share_ctx, gw);
  } else if constexpr (WMAX >= 16) {
    if (W <= 8) enumerate_task<8>(P, m, L, nL, nC, next_id, sumR, nR, ctr, t0, share_ctx, gw);
    else enumerate_task<16>(P, m, L, nL, nC, next_id, sumR, nR, ctr, t0, share_ctx, gw);
  }
}

}  // namespace mbe
