/*
 * gmbe.h — C ABI of the MI355X maximal-biclique-enumeration engine.
 *
 * Every entry point takes plain pointers and sizes; no HIP or torch types
 * cross this boundary.  Return codes: 0 = ok, negative = error (see
 * GMBE_ERR_*); the message of the last error on the calling thread is
 * returned by gmbe_last_error().  The library never calls exit(), and no
 * C++ exception leaves an entry point: a failed host allocation returns
 * GMBE_ERR_NOMEM, anything unexpected GMBE_ERR_INTERNAL.
 *
 * The reference's surface is a C++ class API used from src/main.cu:
 *   Option{uvtrans, order, fast_mode}                 src/BiGraph.h:13-22
 *   CSRBiGraph::Read / CopyToGpu / Reset              src/BiGraph.h:23-38, src/BiGraph.cu:13-278,413-454
 *   BicliqueFinder::Execute / PrintResult             src/BicliqueFinder.h:18-30, src/BicliqueFinder.cu:9-16
 *   IterFinderGpu6(graph, bound_height, bound_size)   src/IterFinderGpu.h:77-90, src/IterFinderGpu.cu:1056-1119
 * Each declaration below names the reference interface it replaces.  The C++
 * drop-in classes with the reference's names live in gmbe_amd/csrc/gmbe.hpp
 * and are built on these functions.
 *
 * Ownership: the caller owns graphs; a finder borrows its graph (it must
 * outlive the finder) and owns its device buffers.  A finder is used by one
 * host thread at a time; distinct finders may run on distinct threads.
 */
#ifndef GMBE_H_
#define GMBE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GMBE_OK 0
#define GMBE_ERR_ARG (-1)         /* invalid argument                       */
#define GMBE_ERR_IO (-2)          /* file cannot be opened / written        */
#define GMBE_ERR_FORMAT (-3)      /* malformed .adj input                   */
#define GMBE_ERR_CAPACITY (-4)    /* a size exceeds a documented limit      */
#define GMBE_ERR_HIP (-5)         /* HIP runtime error or no GPU            */
#define GMBE_ERR_NOMEM (-6)       /* host or device allocation failed       */
#define GMBE_ERR_REJECTED (-7)    /* serving admission rejected a request   */
#define GMBE_ERR_UNSUPPORTED (-8) /* option not supported on this build     */
#define GMBE_ERR_STATE (-9)       /* call out of order (e.g. list before run) */
#define GMBE_ERR_DEADLINE (-10)   /* deadline_ms passed; counts are partial  */
#define GMBE_ERR_INTERNAL (-11)   /* unexpected host exception (a bug)       */

typedef struct gmbe_graph gmbe_graph;
typedef struct gmbe_finder gmbe_finder;
typedef struct gmbe_batch gmbe_batch;
typedef struct gmbe_server gmbe_server;

/* == Option (src/BiGraph.h:13-22); defaults uvtrans=0, order=1, fast_mode=0. */
typedef struct {
  int32_t uvtrans;   /* 0 keep sides; 1 swap if |U|<|V|; 2 swap if |U|>|V|  */
  int32_t order;     /* 0 none; 1 degree asc; 2 degree desc; 3 unilateral   */
  int32_t fast_mode; /* 1 = skip the 2-hop statistics pass                  */
} gmbe_option;

/* Result of one enumeration.  Counts are 64-bit (the reference's int32
 * maximal_nodes_ wraps at 2^31, src/BicliqueFinder.h:28).  The digest is the
 * order-independent hash defined in gmbe_hash_biclique(). */
typedef struct {
  uint64_t maximal;     /* maximal bicliques                                 */
  uint64_t non_maximal; /* diagnostic: rejected task roots + rejected pushes */
  uint64_t hash_sum;    /* sum of per-biclique hashes mod 2^64               */
  uint64_t hash_xor;    /* xor of per-biclique hashes                        */
  uint64_t nodes;       /* enumeration-tree nodes evaluated (roots + pushes) */
  uint64_t tasks;       /* tasks run (top-level vertices + split children)   */
  uint64_t splits;      /* tasks that were split into children               */
  uint64_t overflow;    /* tasks that exceeded a capacity (result invalid)   */
  uint64_t alg_bytes;   /* algorithmic bytes: 4 (deg(x) + |L_parent| + |C|)
                           per push, 4 (deg(S) + sum deg(L)) per task root   */
  double exe_time_s;    /* Execute start -> count on host (exe_time_)        */
  double kernel_time_s; /* enumeration kernel time from HIP events           */
} gmbe_result;

/* ---- graphs ------------------------------------------------------------ */

/* Default Option values (src/BiGraph.h:17-21). */
void gmbe_option_default(gmbe_option* opt);

/* Replaces CSRBiGraph::Read(const char*, Option) (src/BiGraph.cu:13-278).
 * Line i of the file is V vertex i; its tokens are U ids; |U| = max id + 1.
 * Lines are sorted and de-duplicated on load. */
int gmbe_graph_read_adj(const char* path, const gmbe_option* opt,
                        gmbe_graph** out);

/* Same transform pipeline as gmbe_graph_read_adj, from an in-memory CSR in
 * file orientation: row v (0 <= v < nV) lists U ids in [0, nU).  Rows need
 * not be sorted.  Replaces the CSR build at src/BiGraph.cu:251-277. */
int gmbe_graph_from_csr(uint32_t nU, uint32_t nV, const int64_t* v_off,
                        const uint32_t* v_adj, const gmbe_option* opt,
                        gmbe_graph** out);

/* Apply an Option to an already loaded graph (the result keeps original ids
 * of the source graph).  Used to derive the working graph of a generated one. */
int gmbe_graph_derive(const gmbe_graph* src, const gmbe_option* opt,
                      gmbe_graph** out);

/* Statistics in internal orientation: out = {U_size, V_size, edge_size,
 * U_degree, V_degree, U_2degree, V_2degree, swapped}.  The 2-hop degrees are
 * -1 when the graph was loaded with fast_mode (src/BiGraph.h:30). */
int gmbe_graph_stats(const gmbe_graph* g, int64_t out[8]);

/* Borrowed pointers to the host CSR (internal orientation; lists ascending):
 * V->U (row_offset/column_indices) and U->V (rev_row_offset/rev_column_indices),
 * src/BiGraph.h:32-36; plus internal->original id maps of each side.  Any
 * pointer argument may be NULL. */
int gmbe_graph_csr(const gmbe_graph* g, const int64_t** v_off,
                   const uint32_t** v_adj, const int64_t** u_off,
                   const uint32_t** u_adj, const uint32_t** u_orig,
                   const uint32_t** v_orig);

/* Write the graph in .adj format (one line per internal V vertex). */
int gmbe_graph_write_adj(const gmbe_graph* g, const char* path);

/* Write the graph as an edge list in its original orientation and ids: a
 * first line "rows cols" (original V count, original U count), then one line
 * "v u" per edge, rows ascending and each row's columns ascending.  Replaces
 * the .adj -> .graph converter of the reference (preprocess/adj_to_graph.cpp:
 * 31-39); unlike it, rows are written as loaded (sorted, without duplicates)
 * and the column count is the U side's size. */
int gmbe_graph_write_edges(const gmbe_graph* g, const char* path);

/* Replaces CSRBiGraph::Reset() of a host graph (src/BiGraph.cu:433-454). */
void gmbe_graph_destroy(gmbe_graph* g);

/* ---- seeded synthetic workloads -------------------------------------- */

/* Generators of the benchmark graphs.  Output is a graph in file orientation
 * (no side swap, no reordering; original id == internal id).
 *   kind 1: uniform bipartite, nU x nV, m distinct edges
 *   kind 2: Chung-Lu power law, exponents gamma_u/gamma_v, degree cap
 *   kind 3: kind 2 plus `blocks` planted a x b near-bicliques, a,b uniform in
 *           [block_min, block_max], with round(drop*a*b) block edges removed
 * Every vertex gets degree >= 1 before the remaining edges are drawn. */
typedef struct {
  int32_t kind;
  uint32_t nU, nV;
  uint64_t m;
  double gamma_u, gamma_v;
  uint32_t cap;
  uint32_t blocks, block_min, block_max;
  double drop;
  uint64_t seed;
} gmbe_gen_params;

int gmbe_generate(const gmbe_gen_params* p, gmbe_graph** out);

/* One graph of the batched workload: per-side sizes log-uniform in
 * [lo, hi], mean degree uniform in [dlo, dhi], Chung-Lu exponent gamma. */
int gmbe_generate_small(uint64_t seed, uint32_t lo, uint32_t hi, double dlo,
                        double dhi, double gamma, gmbe_graph** out);

/* 2-hop ego subgraph of original-U vertex u of a file-orientation graph:
 * V' = N(u), all edges of V' in ascending V' id until max_edges is reached.
 * The result is in file orientation with ids relabelled densely (ascending). */
int gmbe_ego_subgraph(const gmbe_graph* g, uint32_t u, uint64_t max_edges,
                      gmbe_graph** out);

/* ---- hashing ------------------------------------------------------------ */

/* Hash of one biclique (A = its vertices on the original U side, B = on the
 * original V side, original ids, any order):
 *   key(side, x) = mix64((x << 1) | side)            side 0 = U, 1 = V
 *   SA = sum key(0, a), SB = sum key(1, b)   (mod 2^64)
 *   h  = mix64(SA ^ rotl64(SB, 31))
 * mix64 is the splitmix64 output function (Steele, Lea, Flood 2014).  The
 * digest of an enumeration is (count, sum h mod 2^64, xor h). */
uint64_t gmbe_hash_biclique(const uint32_t* A, size_t nA, const uint32_t* B,
                            size_t nB);

/* ---- the hot path: plain finder --------------------------------------- */

/* Replaces `new IterFinderGpu6(graph, bound_height, bound_size)`
 * (src/IterFinderGpu.h:77-90, src/IterFinderGpu.cu:1056-1085) and, for
 * sel 0/1, IterFinderGpu / IterFinderGpu2 (src/main.cu:136-141): the sel value
 * picks the scheduler (0: one wave per top-level vertex, no splitting;
 * 1: same with larger per-wave prefixes; 2: load-aware splitting, the hot
 * path).  sel 3 (multi-GPU, IterFinderGpu7) returns GMBE_ERR_UNSUPPORTED
 * unless ngpus == 1.  Copies the graph to the device (CopyToGpu,
 * src/BiGraph.cu:413-431). */
int gmbe_finder_create(const gmbe_graph* g, int sel, int bound_height,
                       int bound_size, int ngpus, gmbe_finder** out);

/* Tuning knobs, set before execute (the reference's compile-time macros,
 * src/CMakeLists.txt:13-39, src/GpuUtil.cuh:9-24, src/IterFinderGpu.cu:912-917):
 *   "waves_per_block" (2, 4 or 8), "blocks", "queue_log2" (0..30: 2^queue_log2
 *   queue entries over 32 shards, at least two per shard), "list_cap" (u32
 *   words of list output; 0 = off), "prune" (0/1, PRUNE_EN), "max_prefix"
 *   (2..8), "bound_height", "bound_size".
 * Scheduler and memory knobs (no counterpart): "lds_kb" (LDS list arena per
 *   wave: 0, 3, 7 or 16; default 7), "arena_kb" and "ctx_kb" (per wave;
 *   unless set they scale with |V|: 4-16 MB and 256 KB-1 MB), "ctx_slots",
 *   "group", "share", "share_work", "share_queue", "merge_min",
 *   "merge_branch", "late_share", "profile", "dom_k" (0..16 dominators per
 *   list reduction; 0 = off), "dom_skip" (0/1: mark pushes inside a
 *   dominator non-maximal), "recompute" (0/1: split a large task with no
 *   free context into recomputed children), "count_bytes" (0/1: fill
 *   gmbe_result.alg_bytes).
 * "deadline_ms" (0 = none): every wave stops taking work, and abandons the
 *   subtree it is in, once it has run that long; execute then returns
 *   GMBE_ERR_DEADLINE with the partial counts filled in.
 * Root selection: the run enumerates the tasks of the top-level vertices whose
 *   processing rank r satisfies root_begin <= r < root_end (-1 = all) and
 *   r % root_stride == root_phase; disjoint selections partition the result
 *   (counts and hash sums add, hash xors xor). */
int gmbe_finder_set(gmbe_finder* f, const char* key, int64_t value);

/* Replaces IterFinderGpu6::Execute (src/IterFinderGpu.cu:1095-1119):
 * synchronous; exe_time_s is measured like exe_time_. */
int gmbe_finder_execute(gmbe_finder* f, gmbe_result* res);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Runs the -s 2 enumeration of `g` like
 * gmbe_finder_create + gmbe_finder_execute with the default knobs, but with
 * every device buffer (graph copy, control block, queue, per-wave slabs)
 * carved out of one static device array of 128 MiB that the library holds on
 * each GPU it runs on until the process ends: no hipMalloc, hipFree or other
 * allocating call.  The grid is sized to what fits the array (at most 256
 * waves).  A graph whose copy and one block of slabs do not fit returns
 * GMBE_ERR_CAPACITY before anything runs; a run whose lists outgrow the static
 * arena returns GMBE_ERR_CAPACITY with the partial counts.  Synchronous, on
 * the null stream.  Every call on a GPU shares its array, so calls that
 * overlap on one GPU (from several host threads) race: serialize them. */
int gmbe_finder_execute_static(const gmbe_graph* g, int bound_height, int bound_size,
                               gmbe_result* res);

/* List mode (no reference counterpart: the reference never emits pairs).
 * After an execute with list_cap > 0, copies records
 *   [nA, nB, A[0..nA), B[0..nB)]   (original ids, A on the original U side)
 * into buf.  *used receives the number of words required; returns
 * GMBE_ERR_CAPACITY if cap is too small or the device buffer overflowed. */
int gmbe_finder_list(gmbe_finder* f, uint32_t* buf, size_t cap, size_t* used);

/* Per-block exit times of the last execute with the knob "exit_stamps" = 1:
 * seconds from the earliest block start to each block's end (the -p output
 * of the reference, "SM exit" per SM, src/IterFinderGpu.cu:1038-1041).
 * *n receives the block count; up to cap values are copied into out. */
int gmbe_finder_exit_times(const gmbe_finder* f, double* out, size_t cap, size_t* n);

/* Profile of the last execute (no reference counterpart; per-block exit
 * times are gmbe_finder_exit_times above).  Copies up to n
 * counters: [0] waiting for work, [1] prefix intersections, [2] 2-hop edge
 * pass, [3] extraction + canonicity, [4] candidate-list build, [5] stack
 * machine, [6] split pushes (cycles summed over waves, filled when the knob
 * "profile" is 1); [7] edges of the 2-hop passes, [8] list entries scanned by
 * pushes; [9..11] overflows by cause (list arena, |L| too wide, deferred
 * buffer); [12] list entries and [13] pushes in one-word-bitset lists;
 * [14] cycles of the final wait (a wave with nothing left to do);
 * [15..18] cycles inside the stack machine spent in push passes, twin
 * merges, bitset re-indexing and context publishing; [19] contexts
 * published, [20] twin merges; [21] subtrees abandoned at the deadline;
 * [22] frames handed over by late sharing; [23] cycles spent reducing
 * lists, [24] excluded entries dropped, [25] pushes marked non-maximal;
 * [26] bytes of the bitset lists that pushes read and write plus the CSR
 * bytes of task roots (with the knob "count_bytes");
 * [27] main kernel time and [28] wide-bitset pass time of the last execute
 * (ns, HIP events on the finder's stream), [29] tasks handed to the wide
 * pass; [30] nodes, [31] maximal bicliques, [32] algorithmic bytes and
 * [33] list bytes (both with count_bytes) of the main kernel alone (the rest
 * are the wide pass's).  The cycle counters ([0..6], [14..18], [23]) and
 * [8], [12], [13], [24], [25] are filled when the knob "profile" is 1
 * (else 0); the others always.  Returns the number of counters available. */
int gmbe_finder_stats(gmbe_finder* f, uint64_t* out, size_t n);

/* Replaces the finder destructor chain (src/IterFinderGpu.cu:1087-1093). */
void gmbe_finder_destroy(gmbe_finder* f);

/* ---- batched form ------------------------------------------------------ */

/* Many independent graphs (working form, e.g. from gmbe_graph_derive with
 * -t 1 -o 1 -f) enumerated over a concatenated device CSR with a per-graph
 * table.  Graphs fall in four size classes by edge count ([0,2k), [2k,8k),
 * [8k,32k), [32k,inf)); each class has its own kernel variant and split
 * thresholds, and a run is one persistent launch pair (main and wide pass)
 * per class over that class's graphs.  The run is captured once in a HIP
 * graph and replayed by every gmbe_batch_execute.  No reference counterpart
 * (the reference runs one graph per process, src/main.cu:126-147).  The
 * graphs are borrowed and must outlive the batch. */
int gmbe_batch_create(const gmbe_graph* const* graphs, int n, int bound_height,
                      int bound_size, gmbe_batch** out);
/* Knobs: "waves_per_block" (2, 4, 8; built with lds_kb 7, any of 0/3/7/16,
 * and 7 respectively: setting it moves lds_kb to a built value), "lds_kb"
 * (LDS list arena per wave, KB) -- for every class, or for class c with
 * "waves_per_block.<c>", "lds_kb.<c>"; "bound_height.<c>", "bound_size.<c>"
 * (class c), "prune", "max_prefix", "group", "share", "split", "share_work",
 * "merge_min", "merge_branch", "late_share", "arena_kb", "ctx_kb",
 * "ctx_slots", "dom_k", "dom_skip", "recompute", "count_bytes";
 * "per_graph_launch" = 1 runs the graphs one launch (and one host wait) at a
 * time with their class's kernel, the looped baseline;
 * "autotune" = 1 measures, for each class on a sample of its graphs (8 at
 * most), the kernel variants and then the threshold candidates, and keeps
 * the fastest (synchronous); "verbose" = 1 makes it print one line per trial
 * on stderr. */
int gmbe_batch_set(gmbe_batch* b, const char* key, int64_t value);
/* Read back a knob or a choice of the autotuner: "waves_per_block",
 * "lds_kb" (the defaults), "waves_per_block.<c>", "lds_kb.<c>",
 * "n_classes", "bound_height.<c>", "bound_size.<c>", "class.<i>"
 * (size class of graph i), "kernel_time_us", "wall_time_us" (last execute). */
int gmbe_batch_get(const gmbe_batch* b, const char* key, int64_t* value);
/* per_graph may be NULL; otherwise it receives n results in input order
 * (times zero).  total gets the sums (hash xor combined), the wall time of
 * the run (graph launch to results on the host) and its event time. */
int gmbe_batch_execute(gmbe_batch* b, gmbe_result* per_graph,
                       gmbe_result* total);
void gmbe_batch_destroy(gmbe_batch* b);

/* ---- serving loop ------------------------------------------------------ */

typedef struct {
  uint64_t ticket;      /* id returned by submit                      */
  int32_t status;       /* 0 ok, negative error                        */
  uint64_t maximal, hash_sum, hash_xor;
  double t_submit_s, t_done_s; /* monotonic clock of submit / completion */
} gmbe_serve_done;

/* A server keeps `resident` (file orientation; may be NULL) and runs requests
 * with continuous batching on three lanes, each with its own stream and
 * device buffers: whenever a lane is free, the waiting requests are admitted
 * in arrival order while `slots` (requests per launch) and the lane's share
 * of `mem_budget` (bytes of request data: CSR both ways, keys, table) allow,
 * and launched as one multi-graph run that overlaps the runs of the other
 * lanes.  A request that can never fit is rejected (GMBE_ERR_REJECTED) at
 * submit.  No reference counterpart. */
int gmbe_serve_create(const gmbe_graph* resident, int slots,
                      uint64_t mem_budget, int bound_height, int bound_size,
                      gmbe_server** out);
/* Knobs: "request_ms" (default 0 = none): deadline of each launch; requests
 * of a launch that reaches it finish with status GMBE_ERR_DEADLINE and
 * partial counts, the others with their exact result.  "stack_limit_kb"
 * (default and maximum: the lanes' per-wave list arena, 4096): a request is
 * admitted only if the list memory its largest task root surely needs (one
 * entry per 2-hop candidate, counted from its neighbours' degrees) fits;
 * otherwise submit returns GMBE_ERR_REJECTED. */
int gmbe_serve_set(gmbe_server* s, const char* key, int64_t value);
/* Submit a stand-alone graph (file orientation; copied and brought to the
 * -t 1 -o 1 -f working form on the calling thread). */
int gmbe_serve_submit_graph(gmbe_server* s, const gmbe_graph* g,
                            uint64_t* ticket);
/* Submit the 2-hop ego subgraph of original-U vertex u of the resident graph
 * (gmbe_ego_subgraph with max_edges). */
int gmbe_serve_submit_ego(gmbe_server* s, uint32_t u, uint64_t max_edges,
                          uint64_t* ticket);
/* Collect up to max finished requests; with wait != 0, block until one is
 * finished or none is outstanding. */
int gmbe_serve_poll(gmbe_server* s, gmbe_serve_done* out, int max, int wait,
                    int* n_out);
/* Server statistics, up to n values: [0] launches, [1] requests launched,
 * [2] microseconds the server thread spent packing and uploading, [3]
 * microseconds of launches on the GPU (events, summed over lanes), [4]
 * requests waiting.  Returns the number of values written. */
int gmbe_serve_stats(gmbe_server* s, uint64_t* out, int n);
/* Finishes the requests on the GPU, drops those still waiting.  Submit,
 * poll, set and stats may be called from several threads at once; destroy
 * is called once, after every other call on the server has returned (a
 * poll still blocked in it would wait on freed state). */
void gmbe_serve_destroy(gmbe_server* s);

/* ---- misc ---------------------------------------------------------------- */

/* Thread-local message of the last failing call on this thread. */
const char* gmbe_last_error(void);

/* Device name / CU count of device 0 into buf; GMBE_ERR_HIP without a GPU. */
int gmbe_device_info(char* buf, size_t cap, int* cu_count);

#ifdef __cplusplus
}
#endif

#endif /* GMBE_H_ */
