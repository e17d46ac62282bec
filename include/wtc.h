/* wtc.h -- C ABI of the MI355X-native wedge-parallel triangle counter.
 *
 * The reference (jeffreyspaan/wedge-parallel-triangle-counting) has no library
 * interface: its seam is the internal call tc_GPU() in tc.cu, fed by the C++
 * reader and preprocessing in the same file.  This header exports those seams
 * as plain C functions (plain pointers and sizes, status codes instead of
 * exit()), plus the device-resident, batched and served paths that the
 * reference does not have.
 *
 * Conventions
 *   - Every function returns WTC_OK (0) or a negative WTC_ERR_* code and never
 *     exits the process; wtc_strerror() turns a code into text.
 *   - Host graphs are owned by the caller and only read during a call.
 *   - Device memory is owned by the library (contexts, device graphs, batches).
 *   - One wtc_ctx per host thread; a context is not safe to share without a
 *     lock.  The GPU is device `device` of the HIP runtime.
 */
#ifndef WTC_H
#define WTC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define WTC_OK 0
#define WTC_ERR_ARG (-1)       /* bad argument (e.g. spread 0, a % 32 != 0)   */
#define WTC_ERR_IO (-2)        /* cannot open or read the graph file          */
#define WTC_ERR_FORMAT (-3)    /* no edge line / vertex id range overflow     */
#define WTC_ERR_ZERO_ID (-4)   /* id 0 in a 1-indexed file (tc.cu:599-602)    */
#define WTC_ERR_NOMEM (-5)     /* host allocation failed                      */
#define WTC_ERR_HIP (-6)       /* HIP runtime error (device alloc, launch...) */
#define WTC_ERR_RANGE (-7)     /* graph too large for the 32-bit path         */
#define WTC_ERR_NODEV (-8)     /* no HIP device / kernels not loadable        */
#define WTC_ERR_ADMIT (-9)     /* serving: request exceeds the HBM budget     */
#define WTC_ERR_CSR (-10)      /* CSR fails validation (unsorted, id >= n...) */
#define WTC_ERR_BUSY (-11)     /* serving: queue full / server stopped        */

const char *wtc_strerror(int status);
/* Text of the last HIP failure on this thread (empty if none). */
const char *wtc_last_error_detail(void);

/* ------------------------------------------------------------------- types */

/* GRAPH_TYPE, tc.cu:58-63: oriented CSR with 32-bit offsets. */
typedef struct {
    uint32_t numVertices;
    uint32_t numEdges;
    uint32_t *rowPtr; /* numVertices + 1 */
    uint32_t *colInd; /* numEdges        */
} wtc_graph;

/* BIG_GRAPH_TYPE, experiments/tc_big_graphs.cu:58-63: 64-bit offsets. */
typedef struct {
    uint32_t numVertices;
    uint64_t numEdges;
    uint64_t *rowPtr; /* numVertices + 1 */
    uint32_t *colInd; /* numEdges        */
} wtc_big_graph;

/* GPU_time, tc.cu:76-79: milliseconds, accumulated with += by each call. */
typedef struct {
    double copy;
    double exec;
} wtc_gpu_time;

/* preprocess_t, tc.cu:56.  CPU runs the host builders; GPU and
 * GPU_CONSTRAINED run the sorts on GPU 0 (rocPRIM radix sorts; there is one
 * GPU implementation for both) and return WTC_ERR_NODEV without a GPU.  All
 * three yield the same CSR. */
#define WTC_PREPROCESS_CPU 0
#define WTC_PREPROCESS_GPU 1
#define WTC_PREPROCESS_GPU_CONSTRAINED 2

/* ------------------------------------------------------- host: graph input */

/* read_graph, tc.cu:561-685 (64-bit: experiments/tc_big_graphs.cu:567-691).
 * Edge list (-e) or Matrix Market (-m) text; drops self-loops, symmetrises,
 * sorts, deduplicates; n = largest id + 1.  *out is freed with
 * wtc_free_graph. */
int wtc_read_graph(const char *filename, int matrix_market, int zero_indexed, int preprocess_style,
                   wtc_big_graph **out);

/* The same construction from npairs (src,dst) pairs in memory (interleaved),
 * exactly as if they were the lines of an edge-list file read with -z. */
int wtc_graph_from_pairs(const uint32_t *pairs, uint64_t npairs, wtc_big_graph **out);

/* preprocess, tc.cu:687-787: stable ascending degree order, relabel, keep
 * v->w iff rank(w) > rank(v), rows ascending.  numEdges halves. */
int wtc_preprocess(const wtc_big_graph *original, int preprocess_style, wtc_big_graph **out);

/* free_graph, tc.cu:789-793. */
void wtc_free_graph(wtc_big_graph *g);

/* Binary CSR files (no reference counterpart: the fast-ingestion path).
 * wtc_write_csr stores g with a caller-chosen 64-bit tag (for example a key
 * of the text file and flags it was built from) and a checksum, through a
 * temporary file renamed into place.  wtc_read_csr maps the file, checks
 * size, checksum and structure (offsets non-decreasing from 0 to m, ids < n)
 * and returns WTC_ERR_FORMAT on any mismatch, WTC_ERR_IO if it cannot be
 * opened.  tag may be NULL. */
int wtc_write_csr(const char *path, const wtc_big_graph *g, uint64_t tag);
int wtc_read_csr(const char *path, uint64_t *tag, wtc_big_graph **out);

/* 32-bit view for the tc.cu path; fails with WTC_ERR_RANGE if numEdges does
 * not fit 32 bits.  Free with wtc_free_graph32. */
int wtc_narrow_graph(const wtc_big_graph *g, wtc_graph **out);
void wtc_free_graph32(wtc_graph *g);

/* Checks an oriented CSR: rowPtr monotone, rows strictly ascending, every
 * id < n and greater than its row (the orientation invariant; compare
 * experiments/edge_get.cu:311-316). */
int wtc_validate_oriented(const wtc_big_graph *g);

/* W = sum_v C(d+(v),2), the wedgeSum_total of tc.cu:361. */
uint64_t wtc_wedge_total(const wtc_big_graph *g);

/* Algorithmic bytes of one count over the oriented CSR with offset_bytes-wide
 * row offsets: b_alg = o(n+1) + 4m + 8(n+1) + sum over wedges (v;w,u) of
 * [4 + 4*ceil(log2(d+(w)+1))] (the u read plus a sorted-list probe of
 * adj+(w), one word per binary-search level); b_min = o(n+1) + 4m + 8(n+1)
 * + 8W.  Independent of spread, bitmap and tiling. */
void wtc_alg_bytes(const wtc_big_graph *g, uint32_t offset_bytes, double *b_alg, double *b_min);

/* ------------------------------------------------------ host: generators
 * Seeded, deterministic for any thread count.  Each returns the symmetric,
 * deduplicated CSR that read_graph -z would build from the generated edge
 * list (pass it through wtc_preprocess to orient).  The *_pairs variants
 * return the raw edge list (free with wtc_free_pairs). */

/* R-MAT: edge_factor * 2^scale pairs, quadrant probabilities a,b,c (d=1-a-b-c). */
int wtc_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                 wtc_big_graph **out);
int wtc_gen_rmat_pairs(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                       uint32_t **pairs, uint64_t *npairs);
/* Chung-Lu: n vertices, expected degrees ~ power law (exponent, on
 * [wmin, wmax]) drawn by inverse CDF; endpoints drawn proportional to weight
 * until target_edges distinct undirected non-loop edges exist. */
int wtc_gen_chung_lu(uint32_t n, double exponent, double wmin, double wmax, uint64_t target_edges,
                     uint64_t seed, wtc_big_graph **out);
/* Erdos-Renyi G(n, p). */
int wtc_gen_erdos_renyi(uint32_t n, double p, uint64_t seed, wtc_big_graph **out);
/* Barabasi-Albert: each new vertex attaches to m0 distinct earlier vertices. */
int wtc_gen_barabasi_albert(uint32_t n, uint32_t m0, uint64_t seed, wtc_big_graph **out);
void wtc_free_pairs(uint32_t *pairs);

/* ------------------------------------------------ GPU: reference-style call */

/* tc_GPU, tc.cu:315-441: allocates device buffers, copies the oriented CSR,
 * computes the wedge prefix, search tree and top-a bitmap on the device,
 * counts, copies the count back and frees everything, every call.
 * t->copy += alloc+H2D ms, t->exec += device compute ms (both accumulated,
 * like GPU_time).  adjacency_matrix_len must be a multiple of 32. */
int wtc_tc_gpu(const wtc_graph *graph, uint32_t spread, uint32_t adjacency_matrix_len, wtc_gpu_time *t,
               uint64_t *triangles);
/* tc_GPU of experiments/tc_big_graphs.cu:315-447 (64-bit offsets). */
int wtc_tc_gpu_big(const wtc_big_graph *graph, uint32_t spread, uint32_t adjacency_matrix_len,
                   wtc_gpu_time *t, uint64_t *triangles);
/* tc_arrow_GPU, experiments/tc_arrow.cu:311-441: the arrow-wedge count, per
 * call allocating, copying, counting and freeing like wtc_tc_gpu.  Arrow
 * wedges are the paths v->w->u; u is probed in adj+(v) (bitmap over the
 * top-a ranks, else the sorted list).  Same triangles as wtc_tc_gpu. */
int wtc_tc_arrow_gpu(const wtc_graph *graph, uint32_t spread, uint32_t adjacency_matrix_len, wtc_gpu_time *t,
                     uint64_t *triangles);
/* tc_arrow_GPU of experiments/tc_big_graphs_arrow.cu:334 (64-bit offsets). */
int wtc_tc_arrow_gpu_big(const wtc_big_graph *graph, uint32_t spread, uint32_t adjacency_matrix_len,
                         wtc_gpu_time *t, uint64_t *triangles);

/* wedge_style_t of experiments/tc_vertex.cu:57 (the -w flag). */
#define WTC_WEDGE_OUTGOING 0 /* u after w in adj+(v), probed in adj+(w) */
#define WTC_WEDGE_ARROW 1    /* u in adj+(w), probed in adj+(v)         */
#define WTC_WEDGE_MIXED 2    /* the shorter list walked, the other probed */
#define WTC_PARALLEL_VERTEX 0
#define WTC_PARALLEL_EDGE 1
/* tc_vertex_GPU, experiments/tc_vertex.cu:287 (one lane per vertex), and its
 * 64-bit twin experiments/tc_big_graphs_vertex.cu; per call like wtc_tc_gpu. */
int wtc_tc_vertex_gpu(const wtc_graph *graph, uint32_t adjacency_matrix_len, int style, wtc_gpu_time *t,
                      uint64_t *triangles);
int wtc_tc_vertex_gpu_big(const wtc_big_graph *graph, uint32_t adjacency_matrix_len, int style,
                          wtc_gpu_time *t, uint64_t *triangles);
/* tc_edge_GPU, experiments/tc_edge.cu:374 (one lane per oriented edge), and
 * its 64-bit twin experiments/tc_big_graphs_edge.cu. */
int wtc_tc_edge_gpu(const wtc_graph *graph, uint32_t adjacency_matrix_len, int style, wtc_gpu_time *t,
                    uint64_t *triangles);
int wtc_tc_edge_gpu_big(const wtc_big_graph *graph, uint32_t adjacency_matrix_len, int style,
                        wtc_gpu_time *t, uint64_t *triangles);

/* --------------------------------------------- GPU: device-resident graphs */

typedef struct wtc_ctx wtc_ctx;       /* HIP stream + reusable workspace */
typedef struct wtc_dgraph wtc_dgraph; /* oriented CSR resident in HBM    */

typedef struct {
    double prefix_ms;  /* wedge prefix scan + setup (split count: its two
                          prefixes and range map)                             */
    double bitmap_ms;  /* bitmap clear + build (split count: the counting
                          sort of the in-edges by middle vertex and the
                          packing of the top tails' ids)                      */
    double count_ms;   /* counting kernel(s) (split count: row + slab kernels) */
    double total_ms;   /* first launch to count back on the host             */
    uint64_t wedges;   /* W counted                                          */
    uint32_t blocks;   /* counting-kernel grid                               */
    uint32_t launches; /* counting-kernel launches                           */
    uint32_t small_graphs; /* graphs counted by the small-graph kernel       */
    uint32_t split;    /* 1: counted by the split count                      */
    double rows_ms;    /* split count: its row kernel alone                  */
    uint64_t top_wedges; /* split count: wedges probed in LDS (w in the top rows) */
    uint64_t in_edges; /* split count: in-edge tails the row kernel streamed */
    double list_ms;    /* split count: its slab kernel alone (w below the top rows) */
    uint64_t list_in_edges; /* split count: in-edge tails the slab kernel streamed */
} wtc_count_stats;

int wtc_device_count(int *count);
int wtc_ctx_create(int device, wtc_ctx **out);
void wtc_ctx_destroy(wtc_ctx *ctx);
/* Tuning of the counting kernel: probe_ways = independent binary searches
 * each lane advances together (1, 2, 4 or 8; default 4 or $WTC_PROBE_WAYS);
 * lds_chunk = wedges per lane staged in LDS per phase (1..16, default 16);
 * waves_per_block = the counting block's shape, 1, 2, 4 or 8 waves of 64
 * lanes (default 4 or $WTC_COUNT_WAVES).  Counts do not depend on these. */
int wtc_ctx_set_tuning(wtc_ctx *ctx, int probe_ways, uint32_t lds_chunk, int waves_per_block);
/* Counting kernel of a context (both count the same tiles of the same wedge
 * order, tc.cu:173-293, with the same probes):
 *   WTC_KERNEL_WALK   each wave walks its tiles' wedges in steps of 64
 *                     consecutive wedges (long rows: u loaded straight from
 *                     the row; each vertex's last 64 entries: pairs from a
 *                     register window); probe_ways = steps in flight.
 *   WTC_KERNEL_STAGED each lane decodes its spread of s wedges and stages
 *                     them through LDS, as the reference does
 *                     (tc.cu:174-244); lds_chunk applies.
 *   WTC_KERNEL_SPLIT  a single graph with a > 0: every in-edge v -> w is
 *                     grouped by its middle vertex w with a counting sort
 *                     per count; the wedges whose w is among the top-a
 *                     ranks (the reference's bitmap probes, tc.cu:262-272)
 *                     are counted against adj+(w) held as a bitmap row in
 *                     LDS (their tails streamed from a copy of colInd
 *                     packed to 16-21 bits per id), the rest (sorted-list
 *                     probes, tc.cu:273-286) against hash sets of adj+(w)
 *                     in LDS.  No a*a/8-byte bitmap in HBM; about 27 or 43
 *                     bytes per edge of workspace (4- or 8-byte offsets)
 *                     instead.  Batches, a = 0, a row of a bits that does
 *                     not fit a block's LDS and colInd not 16-byte aligned:
 *                     the row walk.  Same count for every wedge range.
 * Default WTC_KERNEL_WALK or $WTC_COUNT_KERNEL. */
#define WTC_KERNEL_WALK 0
#define WTC_KERNEL_STAGED 1
#define WTC_KERNEL_SPLIT 2
int wtc_ctx_set_count_kernel(wtc_ctx *ctx, int kernel);
/* How the counting kernels hand out work: the wedge order is cut into tiles
 * of 64*s wedges, and a wave takes up to tiles_per_grab consecutive tiles
 * per atomic grab (1..65535; 0 = the kernel's default: 4 for the walk, 32
 * for the staged kernel).  deal_to_xcds != 0 deals super-chunks of grabs to
 * the 8 XCDs in turn, each XCD taking from its own counter; 0 gives one
 * stream of grabs for the whole grid.  Default 0 (the kernel's), one stream.
 * A grab covers at least 2048 wedges whatever tiles_per_grab says, and a
 * launch with few tiles takes smaller grabs (about two per wave).  Counts do
 * not depend on these. */
int wtc_ctx_set_work_split(wtc_ctx *ctx, uint32_t tiles_per_grab, int deal_to_xcds);
/* Kernel choice for sets of small graphs (the role of the vertex- and
 * edge-parallel kernels, experiments/tc_vertex.cu:113-274 and
 * experiments/tc_edge.cu:179-352, for small graphs in a batch).  A count over
 * two or more graphs (a batch, a serving batch) in which every graph has
 * 2(n+1) + m <= max_words is done by the small-graph kernel: one block per
 * graph holds the graph's CSR and wedge prefix in LDS.  Any other set is
 * done by the wedge-parallel kernel.  0 turns the small-graph kernel off.
 * max_words may not exceed the device's LDS per block in words
 * (device_max_words; 40960 on MI355X).  Default 20480 or
 * $WTC_SMALL_GRAPH_WORDS.  Counts do not depend on it. */
int wtc_ctx_set_small_graphs(wtc_ctx *ctx, uint32_t max_words);
int wtc_ctx_get_small_graphs(const wtc_ctx *ctx, uint32_t *max_words, uint32_t *device_max_words);

/* Copies an oriented CSR to the device.  offset_bytes 4 or 8 selects the
 * device rowPtr width (0: 4 if numEdges < 2^32 else 8). */
int wtc_dgraph_upload(wtc_ctx *ctx, const wtc_big_graph *g, uint32_t offset_bytes, wtc_dgraph **out);
void wtc_dgraph_free(wtc_dgraph *dg);
uint64_t wtc_dgraph_bytes(const wtc_dgraph *dg);
/* Vertices, oriented edges and row-offset width (4 or 8) of a device graph. */
int wtc_dgraph_info(const wtc_dgraph *dg, uint32_t *n, uint64_t *m, uint32_t *offset_bytes);

/* Builds the oriented CSR on the device from (src,dst) host pairs, as
 * read_graph -z then preprocess would (GPU preprocessing, result identical to
 * the CPU builders); offset_bytes as in wtc_dgraph_upload. */
int wtc_dgraph_from_pairs(wtc_ctx *ctx, const uint32_t *pairs, uint64_t npairs, uint32_t offset_bytes,
                          wtc_dgraph **out);
/* The R-MAT graph of wtc_gen_rmat (same pairs, drawn on the device), oriented
 * on the device. */
int wtc_dgraph_gen_rmat(wtc_ctx *ctx, uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                        uint64_t seed, uint32_t offset_bytes, wtc_dgraph **out);
/* Copies a resident oriented CSR back to the host (free with wtc_free_graph). */
int wtc_dgraph_download(const wtc_dgraph *dg, wtc_big_graph **out);

/* Counts the wedges [w_begin, w_end) (clamped to [0, W); pass 0, UINT64_MAX
 * for all) of a resident graph: the device part of tc_GPU, from the resident
 * CSR to the count on the host.  `st` may be NULL. */
int wtc_dgraph_count(wtc_ctx *ctx, wtc_dgraph *dg, uint32_t spread, uint32_t adjacency_matrix_len,
                     uint64_t w_begin, uint64_t w_end, uint64_t *triangles, wtc_count_stats *st);

/* The arrow-wedge count of a resident graph (the device part of tc_arrow_GPU,
 * experiments/tc_arrow.cu:173-289 and 345-368): arrow-wedge prefix over the
 * edges (rows with d+ < 2 carry none), bitmap, count.  *arrow_wedges (may be
 * NULL) receives the arrow-wedge total; st->wedges holds it too. */
int wtc_dgraph_count_arrow(wtc_ctx *ctx, wtc_dgraph *dg, uint32_t spread, uint32_t adjacency_matrix_len,
                           uint64_t *triangles, uint64_t *arrow_wedges, wtc_count_stats *st);

/* The vertex- or edge-parallel baseline count of a resident graph (the device
 * part of tc_vertex_GPU / tc_edge_GPU, experiments/tc_vertex.cu:113-274,
 * experiments/tc_edge.cu:179-352): parallelism WTC_PARALLEL_*, style
 * WTC_WEDGE_*.  st->bitmap_ms and st->count_ms are filled; st->wedges is 0. */
int wtc_dgraph_count_baseline(wtc_ctx *ctx, wtc_dgraph *dg, int parallelism, int style,
                              uint32_t adjacency_matrix_len, uint64_t *triangles, wtc_count_stats *st);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The count of wtc_dgraph_count over a caller-owned
 * oriented CSR already in device memory (d_row_ptr: n+1 offsets of
 * offset_bytes 4 or 8; d_col_ind: m ids), with no device allocation: every
 * workspace buffer comes from one fixed array of 256 MiB per device that the
 * library holds until the process ends.  The workspace 8(n+1) +
 * 8*ceil(n/4096) + a*a/8 bytes (plus ~34 KB) must fit it, else WTC_ERR_RANGE
 * before anything is enqueued (R-MAT-24 fits with a <= 16384).
 * Kernels only are enqueued on `stream` (a hipStream_t; NULL = the null
 * stream of the current device); *d_triangles (8 bytes of device memory)
 * holds the count once the stream has reached that point.  The CSR is
 * checked on the device first: one that is not an oriented CSR is not
 * counted and *d_triangles becomes UINT64_MAX.  Calls that overlap on one
 * GPU share its array and race: serialise them (one stream, or a lock). */
int wtc_count_noalloc(const void *d_row_ptr, uint32_t offset_bytes, const uint32_t *d_col_ind, uint32_t n,
                      uint64_t m, uint32_t spread, uint32_t adjacency_matrix_len, uint64_t w_begin, uint64_t w_end,
                      uint64_t *d_triangles, void *stream);

/* Diagnostics for parity tests: the device wedge prefix (n+1 entries, the
 * h_wedgeSum of tc.cu:348-361), the 4095-entry search tree (the
 * c_binary_search_cache of tc.cu:364-368, filled for single-graph counts) and
 * the bitmap words (tc.cu:329-346) left by the last count on `ctx`. */
int wtc_ctx_debug_prefix(wtc_ctx *ctx, uint64_t n, uint64_t *wedge_sum, uint64_t *tree);
int wtc_ctx_debug_bitmap(wtc_ctx *ctx, uint64_t words, uint32_t *out);
/* Copies `bytes` (rounded down to 16) between two fresh device buffers `reps`
 * times with a plain streaming kernel; *ms = average per copy.  The byte
 * count is known exactly, so profiler HBM counters read on it calibrate the
 * counters read on the counting kernel. */
int wtc_debug_copy(wtc_ctx *ctx, uint64_t bytes, int reps, double *ms);
/* Runs the device first-wedge decode (tc.cu:209-212) on k host-given
 * (d, i_v) pairs, d >= 2 and i_v < C(d,2), writing (w_i, u_i). */
int wtc_debug_decode(wtc_ctx *ctx, const uint32_t *d, const uint64_t *iv, uint64_t k, uint32_t *w_out,
                     uint32_t *u_out);

/* -------------------------------------------------------- GPU: batching
 * Many graphs counted by one persistent launch.  The graphs are laid out in
 * HBM as one disjoint union (ids shifted); every tile of the wedge space
 * belongs to one graph, and each graph has its own spread and bitmap
 * length.  spread/a arrays of length k; a spread of 0 asks the autotuner. */

typedef struct wtc_batch wtc_batch;

int wtc_batch_upload(wtc_ctx *ctx, const wtc_big_graph *const *graphs, size_t k, wtc_batch **out);
void wtc_batch_free(wtc_batch *b);
/* Counts a resident batch (ctx must be the one it was uploaded with).  The
 * launch sequence is captured into a HIP graph on the batch's own stream and
 * workspace the first time, and replayed as one graph launch while spread,
 * a and the context's tuning stay the same (a change captures anew).  In
 * *st, count_ms is the device time of the whole replay (prefix, bitmap and
 * count); prefix_ms and bitmap_ms are 0. */
int wtc_batch_count(wtc_ctx *ctx, wtc_batch *b, const uint32_t *spread, const uint32_t *adjacency_matrix_len,
                    uint64_t *triangles, wtc_count_stats *st);
/* Host CSRs in, counts out in one call (the batched counterpart of calling
 * wtc_tc_gpu once per graph): each graph is checked and packed into pinned
 * staging, every chunk of $WTC_UPLOAD_CHUNK_BYTES (default 256 MiB) is copied
 * to the device and counted as its own graph set as soon as it lands, on two
 * execution slots of the context, so counting overlaps the upload.  The
 * device buffers stay in the context for the next call.  A graph that is not
 * an oriented CSR fails the whole call with WTC_ERR_CSR. */
int wtc_batch_count_host(wtc_ctx *ctx, const wtc_big_graph *const *graphs, size_t k, const uint32_t *spread,
                         const uint32_t *adjacency_matrix_len, uint64_t *triangles, wtc_count_stats *st);
/* ---------------------------------------------------------- GPU: serving
 * A request loop with continuous batching over device-resident CSRs (no
 * counterpart in the reference).  Graphs are registered once; requests name
 * a graph.  A worker thread admits requests by HBM footprint (resident CSR +
 * per-request wedge prefix and bitmap, carved from a fixed device pool;
 * least-recently-used idle graphs are evicted), uploads missing graphs on
 * their own stream, and starts a batch (one graph set, one launch sequence)
 * whenever one of `slots` execution slots is free, so later requests run
 * while earlier batches are still on the device.  Thread-safe. */
typedef struct wtc_server wtc_server;
typedef struct {
    uint64_t hbm_budget_bytes; /* device pool (default 8 GiB)                     */
    uint32_t max_batch;        /* graphs per batch (default 256)                  */
    uint32_t slots;            /* batches in flight (default 2, at most 8)        */
    uint32_t queue_capacity;   /* queued requests before WTC_ERR_BUSY (65536)     */
    int32_t probe_ways;        /* counting-kernel tuning, 0 = default             */
    uint32_t lds_chunk;        /* counting-kernel tuning, 0 = default             */
    int32_t waves_per_block;   /* counting block: 1, 2, 4 or 8 waves, 0 = 4       */
    int32_t small_graph_words; /* small-graph kernel limit (wtc_ctx_set_small_graphs):
                                  0 = default, -1 = off, else the limit in words */
} wtc_server_config;
typedef struct {
    double queue_ms;      /* submit -> batch start                                */
    double service_ms;    /* batch start -> completion seen                       */
    double latency_ms;    /* submit -> completion seen                            */
    uint32_t batch_graphs;
    uint64_t batch_id;
} wtc_request_stats;
typedef struct {
    uint64_t graphs, submitted, completed, rejected, batches, max_batch_graphs;
    uint64_t uploads, evictions, resident_bytes, resident_graphs, queued;
} wtc_server_stats;

int wtc_server_create(int device, const wtc_server_config *cfg, wtc_server **out);
/* Validates and copies an oriented CSR (numEdges < 2^32) to pinned host
 * memory; the caller's graph may be freed afterwards. */
int wtc_server_register(wtc_server *srv, const wtc_big_graph *oriented, uint64_t *graph_id);
/* WTC_ERR_ADMIT if the request can never fit the pool, WTC_ERR_BUSY if the
 * queue is full. */
int wtc_server_submit(wtc_server *srv, uint64_t graph_id, uint32_t spread, uint32_t adjacency_matrix_len,
                      uint64_t *request_id);
/* Waits (timeout_us < 0: forever; 0: poll) for a request; WTC_ERR_BUSY if not
 * done in time.  A finished request is forgotten after a successful wait. */
int wtc_server_wait(wtc_server *srv, uint64_t request_id, int64_t timeout_us, uint64_t *triangles,
                    wtc_request_stats *st);
int wtc_server_get_stats(wtc_server *srv, wtc_server_stats *st);
/* Finishes queued and running requests, then frees everything. */
void wtc_server_destroy(wtc_server *srv);

/* The autotuner: choices for a graph of n vertices, m oriented edges and W
 * wedges from a per-size-bucket table measured on MI355X (the sweep
 * `bench.py --tune` writes tuning/mi355x.tsv beside the library;
 * $WTC_TUNING_FILE names another; built-in rows otherwise).  Buckets are
 * floor(log2 W) by family (counted alone, or as a member of a batch, where
 * many graphs share one launch) and degree skew (W/m >= 24).  No counterpart
 * in the reference, which leaves s and a to the user (tc.cu:828-881). */
typedef struct wtc_tuned {
    uint32_t spread;               /* s */
    uint32_t adjacency_matrix_len; /* a (0: no bitmap) */
    int32_t kernel;                /* WTC_KERNEL_* */
    int32_t waves_per_block;       /* counting block: 1, 2, 4 or 8 waves */
    int32_t probe_ways;            /* probes (walk: steps) in flight per lane */
    int32_t split_waves;           /* split count's row kernel: 4, 8 or 16 waves */
} wtc_tuned;
int wtc_autotune_shape(uint32_t n, uint64_t m, uint64_t wedges, int batch, wtc_tuned *out);
/* The table's (spread, a) alone: wtc_autotune for a graph counted alone,
 * wtc_autotune_batch for a member of a batch. */
void wtc_autotune(uint32_t n, uint64_t m, uint64_t wedges, uint32_t *spread, uint32_t *adjacency_matrix_len);
void wtc_autotune_batch(uint32_t n, uint64_t m, uint64_t wedges, uint32_t *spread, uint32_t *adjacency_matrix_len);
/* Path of the table in use, or "built-in". */
const char *wtc_tuning_source(void);
/* Applies a tuned shape's kernel, block waves, probe ways and row-kernel
 * waves to a context (spread and a stay per count). */
int wtc_ctx_use_tuned(wtc_ctx *ctx, const wtc_tuned *t);

#ifdef __cplusplus
}
#endif
#endif
