/* oracle.h -- CPU restatement of WeTriC's counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the libwtc library, the
 * tc CLI, the HIP kernels) links, loads or calls this code.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only as
 * the checker.
 *
 * Every function restates one piece of the reference (jeffreyspaan/
 * wedge-parallel-triangle-counting) and cites the file:line it follows.
 * The reference (tc.cu) needs nvcc, CUB and libcudart, none of which exist in
 * this image, so it cannot be compiled here; the restatement is pinned instead
 * by the checks listed in the tests (closed forms, brute force, networkx golden
 * fixtures, exhaustive comparison with the reference's floating-point decode).
 *
 * Graphs use 64-bit row offsets throughout (the BIG_GRAPH_TYPE layout of
 * experiments/tc_big_graphs.cu:58-63); the 32-bit GRAPH_TYPE of tc.cu:58-63
 * is the same data with narrower offsets.
 */
#ifndef WTC_ORACLE_H
#define WTC_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t n;        /* numVertices */
    uint64_t m;        /* numEdges (entries of colInd) */
    uint64_t *rowPtr;  /* n+1 */
    uint32_t *colInd;  /* m */
} orc_graph;

#define ORC_OK 0
#define ORC_ERR_IO (-2)
#define ORC_ERR_FORMAT (-3)
#define ORC_ERR_ZERO_ID (-4)
#define ORC_ERR_NOMEM (-5)
#define ORC_ERR_ARG (-1)

/* tc.cu:30-31: 12 cached levels, 4095 entries. */
#define ORC_TREE_LEVELS 12
#define ORC_TREE_SIZE ((1u << ORC_TREE_LEVELS) - 1u)

/* read_graph, tc.cu:561-685.  Symmetric, deduplicated CSR, self-loops
 * dropped, n = max id + 1.  All -p styles produce the same CSR (the sort key
 * (src,dst) is a total order), so there is one function. */
int orc_read_graph(const char *path, int matrix_market, int zero_indexed, orc_graph *out);

/* The same construction from an in-memory list of (src,dst) pairs, as if the
 * pairs were the lines of an edge-list file read with -z (tc.cu:590-685). */
int orc_graph_from_pairs(const uint32_t *pairs, uint64_t npairs, orc_graph *out);

/* preprocess, tc.cu:687-787: rank = stable ascending sort by degree (ties by
 * id, which is what CUB's stable radix sort gives for the default -p 2),
 * keep v->w iff rank(w) > rank(v), rows ascending. */
int orc_preprocess(const orc_graph *in, orc_graph *out);

void orc_free(orc_graph *g);

/* Host wedge prefix, tc.cu:348-361 (64-bit form of
 * experiments/tc_big_graphs.cu:352-358): wedgeSum[0]=0,
 * wedgeSum[v+1] = wedgeSum[v] + C(d+(v),2).  Returns W = wedgeSum[n]. */
uint64_t orc_wedge_prefix(const orc_graph *g, uint64_t *wedgeSum /* n+1 */);

/* build_binary_search_cache, tc.cu:306-313 (called at tc.cu:367). */
void orc_search_tree(const uint64_t *wedgeSum, uint32_t n, uint64_t *tree /* ORC_TREE_SIZE */);

/* binary_search_closest_ULONG_GPU, tc.cu:117-137: rightmost index in
 * [start,end) whose value is <= target. */
uint32_t orc_search_closest(const uint64_t *list, uint32_t start, uint32_t end, uint64_t target);

/* binary_search_closest_ULONG_constant_GPU, tc.cu:139-170. */
uint32_t orc_search_closest_tree(const uint64_t *tree, const uint64_t *list,
                                 uint32_t start, uint32_t end, uint64_t target);

/* First-wedge decode, tc.cu:209-212, restated with integer-exact arithmetic:
 * local wedge index i_v of a vertex with out-degree d -> (w_i, u_i),
 * 0 <= w_i < u_i < d, row-major over the strict upper triangle. */
void orc_decode(uint64_t d, uint64_t i_v, uint64_t *w_i, uint64_t *u_i);

/* The reference's own floating-point formula, tc.cu:211-212 (64-bit form of
 * experiments/tc_big_graphs.cu:211-212), kept only so tests can compare the
 * integer decode against it. */
void orc_decode_ref_fp(uint64_t d, uint64_t i_v, uint64_t *w_i, uint64_t *u_i);

/* Exhaustive check of both decodes against row-by-row enumeration for every
 * d in [d_lo, d_hi]; 0 if all agree, else the first failing d. */
uint64_t orc_decode_check(uint64_t d_lo, uint64_t d_hi, int fp, uint64_t *bad_i);

/* Merge count restricted to rows [v0, v1) (a bounded sample of the CPU
 * baseline). */
uint64_t orc_tc_merge_range(const orc_graph *g, uint32_t v0, uint32_t v1, int threads);

/* Bitmap over the top-a ranks: index (tc.cu:263, 338) and host build
 * (tc.cu:329-346).  words must hold orc_bitmap_words(a) zeroed words. */
uint64_t orc_bitmap_bits(uint32_t a);
uint64_t orc_bitmap_words(uint32_t a);
uint64_t orc_bitmap_index(uint32_t n, uint32_t a, uint32_t w, uint32_t u);
void orc_bitmap_build(const orc_graph *g, uint32_t a, uint32_t *words);

/* The wedge-parallel count of tc_GPU_kernel (tc.cu:173-293) as launched by
 * tc_GPU (tc.cu:394-414), executed sequentially: blocks of 128 threads, each
 * thread `spread` consecutive wedges (cached search + decode for the first,
 * incremental stepping after), then the cooperative probe over the block's
 * wedges (bitmap for w in the top-a ranks, else linear/binary search).
 * wedge range [w_begin, w_end) as in experiments/tc_big_graphs.cu:394-420.
 * Returns ORC_OK and writes the count. */
int orc_tc_wedge(const orc_graph *g, uint32_t spread, uint32_t a,
                 uint64_t w_begin, uint64_t w_end, uint64_t *count);

/* Arrow-wedge prefix of experiments/tc_arrow.cu:345-368 (vi: m+1 entries;
 * edges of rows with d+ < 2 carry no wedges).  Returns the arrow-wedge total. */
uint64_t orc_arrow_prefix(const orc_graph *g, uint64_t *vi);

/* The arrow-wedge count of tc_arrow_GPU_kernel (experiments/tc_arrow.cu:
 * 173-289): wedges v->w->u, u probed in adj+(v) (bitmap over the top-a
 * ranks, else list search).  Writes the arrow-wedge total and the count. */
int orc_tc_arrow(const orc_graph *g, uint32_t a, uint64_t *arrow_wedges, uint64_t *count);

/* The vertex-/edge-parallel baselines' count in wedge style 0 outgoing,
 * 1 arrow, 2 mixed (experiments/tc_vertex.cu:113-274, tc_edge.cu:179-352). */
int orc_tc_style(const orc_graph *g, int style, uint32_t a, uint64_t *count);

/* Merge-intersection count over the same oriented CSR, with `threads` POSIX
 * threads (<=0: all online cores).  This is the CPU baseline: for each v, for
 * each w in adj+(v), add |adj+(v) after w  intersect  adj+(w)|. */
uint64_t orc_tc_merge(const orc_graph *g, int threads);

/* Algorithmic bytes of the count: B_alg = o(n+1) + 4m + 8(n+1)
 * + sum over wedges (v;w,u) of [4 + 4*ceil(log2(d+(w)+1))], and the floor
 * B_min = o(n+1) + 4m + 8(n+1) + 8W. */
void orc_alg_bytes(const orc_graph *g, uint32_t offset_bytes, double *b_alg, double *b_min);

#ifdef __cplusplus
}
#endif
#endif
