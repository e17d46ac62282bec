/* oracle.c -- CPU restatement of WeTriC's counting path (test infrastructure
 * only; see oracle.h).  Plain C11 + POSIX threads.  Each function cites the
 * reference lines it restates. */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

static uint64_t binom2(uint64_t x) { return (x & 1u) ? x * ((x - 1u) >> 1) : (x >> 1) * (x - 1u); }

void orc_free(orc_graph *g) {
    if (!g) return;
    free(g->rowPtr);
    free(g->colInd);
    g->rowPtr = NULL;
    g->colInd = NULL;
    g->n = 0;
    g->m = 0;
}

/* ------------------------------------------------------------------ reader */

/* In-memory stand-in for a FILE*, so that fgets(line,256) and
 * fscanf("%d %d\n") can be followed exactly (tc.cu:569-620). */
typedef struct {
    const char *buf;
    size_t len, pos;
} orc_stream;

/* fgets(line, 256, f): up to 255 chars, through the first '\n'. */
static int stream_fgets(orc_stream *s, char *line, size_t cap) {
    if (s->pos >= s->len) return 0;
    size_t k = 0;
    while (k + 1 < cap && s->pos < s->len) {
        char c = s->buf[s->pos++];
        line[k++] = c;
        if (c == '\n') break;
    }
    line[k] = 0;
    return 1;
}

static int is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

/* One "%d" conversion: skip white space, optional sign, decimal digits; the
 * value goes through long (saturating) and is then narrowed to 32 bits, as
 * glibc's scanf does for an int target read into UINT_t (tc.cu:588-590). */
static int scan_int(const char *b, size_t len, size_t *pos, uint32_t *out) {
    size_t p = *pos;
    while (p < len && is_ws(b[p])) p++;
    int neg = 0;
    if (p < len && (b[p] == '+' || b[p] == '-')) { neg = b[p] == '-'; p++; }
    if (p >= len || b[p] < '0' || b[p] > '9') { *pos = p; return 0; }
    unsigned long long v = 0;
    int sat = 0;
    while (p < len && b[p] >= '0' && b[p] <= '9') {
        if (!sat) {
            v = v * 10u + (unsigned)(b[p] - '0');
            if (v > 0x7fffffffffffffffULL) sat = 1;
        }
        p++;
    }
    long long sv;
    if (sat) sv = neg ? (long long)(-0x7fffffffffffffffLL - 1) : 0x7fffffffffffffffLL;
    else sv = neg ? -(long long)v : (long long)v;
    *out = (uint32_t)(uint64_t)sv;
    *pos = p;
    return 1;
}

/* "%d %d\n": two conversions, then trailing white space. Returns #converted. */
static int scan_pair(const char *b, size_t len, size_t *pos, uint32_t *v, uint32_t *w) {
    if (!scan_int(b, len, pos, v)) return 0;
    if (!scan_int(b, len, pos, w)) return 1;
    while (*pos < len && is_ws(b[*pos])) (*pos)++;
    return 2;
}

typedef struct {
    uint32_t *p; /* interleaved src,dst */
    uint64_t n, cap;
} pairvec;

static int pv_push(pairvec *pv, uint32_t a, uint32_t b) {
    if (pv->n + 1 > pv->cap) {
        uint64_t nc = pv->cap ? pv->cap * 2 : 4096;
        uint32_t *np = (uint32_t *)realloc(pv->p, nc * 2 * sizeof(uint32_t));
        if (!np) return ORC_ERR_NOMEM;
        pv->p = np;
        pv->cap = nc;
    }
    pv->p[2 * pv->n] = a;
    pv->p[2 * pv->n + 1] = b;
    pv->n++;
    return ORC_OK;
}

static int cmp_pair(const void *x, const void *y) {
    const uint32_t *a = (const uint32_t *)x, *b = (const uint32_t *)y;
    if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
    if (a[1] != b[1]) return a[1] < b[1] ? -1 : 1;
    return 0;
}

/* Sort + dedup + CSR, tc.cu:625-682 (qsort path of tc.cu:646). */
static int csr_from_directed(pairvec *pv, uint32_t max_vertex, orc_graph *out) {
    uint64_t nv = (uint64_t)max_vertex + 1u;
    if (nv > 0xffffffffULL) return ORC_ERR_FORMAT;
    out->n = (uint32_t)nv;
    out->rowPtr = (uint64_t *)calloc(nv + 1, sizeof(uint64_t));
    if (!out->rowPtr) return ORC_ERR_NOMEM;
    qsort(pv->p, pv->n, 2 * sizeof(uint32_t), cmp_pair);
    uint64_t k = 0;
    out->colInd = (uint32_t *)malloc((pv->n ? pv->n : 1) * sizeof(uint32_t));
    if (!out->colInd) return ORC_ERR_NOMEM;
    for (uint64_t i = 0; i < pv->n; i++) {
        if (i > 0 && pv->p[2 * i] == pv->p[2 * i - 2] && pv->p[2 * i + 1] == pv->p[2 * i - 1]) continue;
        out->colInd[k++] = pv->p[2 * i + 1];
        out->rowPtr[pv->p[2 * i] + 1]++;
    }
    for (uint64_t v = 1; v <= nv; v++) out->rowPtr[v] += out->rowPtr[v - 1];
    out->m = k;
    return ORC_OK;
}

/* Per-pair handling of tc.cu:599-619: zero-id check, 1->0 index shift, drop
 * self-loops, push both directions, track the largest id. */
static int add_pair(pairvec *pv, uint32_t v, uint32_t w, int zero_indexed, uint32_t *max_vertex) {
    if (!zero_indexed && (v == 0 || w == 0)) return ORC_ERR_ZERO_ID;
    v -= zero_indexed ? 0u : 1u;
    w -= zero_indexed ? 0u : 1u;
    if (v == w) return ORC_OK;
    uint32_t mx = v > w ? v : w;
    if (mx > *max_vertex) *max_vertex = mx;
    int rc = pv_push(pv, v, w);
    if (rc) return rc;
    return pv_push(pv, w, v);
}

/* read_graph, tc.cu:561-685. */
int orc_read_graph(const char *path, int matrix_market, int zero_indexed, orc_graph *out) {
    memset(out, 0, sizeof(*out));
    FILE *f = fopen(path, "rb");
    if (!f) return ORC_ERR_IO;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = (char *)malloc((size_t)sz + 1);
    if (!buf) { fclose(f); return ORC_ERR_NOMEM; }
    size_t got = fread(buf, 1, (size_t)sz, f);
    fclose(f);
    orc_stream s = {buf, got, 0};
    char line[256];
    /* tc.cu:571-574: skip lines whose first character is not a digit. */
    do {
        if (!stream_fgets(&s, line, sizeof line)) { free(buf); return ORC_ERR_FORMAT; }
    } while (line[0] < '0' || line[0] > '9');
    /* tc.cu:577-579: Matrix Market size line. */
    if (matrix_market && !stream_fgets(&s, line, sizeof line)) { free(buf); return ORC_ERR_FORMAT; }

    pairvec pv = {0};
    uint32_t max_vertex = 0, v, w;
    size_t lp = 0;
    int rc = ORC_OK;
    /* tc.cu:590: the first pair comes from the line buffer ... */
    if (scan_pair(line, strlen(line), &lp, &v, &w) == 2) {
        rc = add_pair(&pv, v, w, zero_indexed, &max_vertex);
        /* ... tc.cu:620: the rest from the stream. */
        while (rc == ORC_OK && scan_pair(s.buf, s.len, &s.pos, &v, &w) == 2)
            rc = add_pair(&pv, v, w, zero_indexed, &max_vertex);
    }
    free(buf);
    if (rc == ORC_OK) rc = csr_from_directed(&pv, max_vertex, out);
    free(pv.p);
    if (rc) orc_free(out);
    return rc;
}

int orc_graph_from_pairs(const uint32_t *pairs, uint64_t npairs, orc_graph *out) {
    memset(out, 0, sizeof(*out));
    pairvec pv = {0};
    uint32_t max_vertex = 0;
    int rc = ORC_OK;
    for (uint64_t i = 0; i < npairs && rc == ORC_OK; i++)
        rc = add_pair(&pv, pairs[2 * i], pairs[2 * i + 1], 1, &max_vertex);
    if (rc == ORC_OK) rc = csr_from_directed(&pv, max_vertex, out);
    free(pv.p);
    if (rc) orc_free(out);
    return rc;
}

/* ------------------------------------------------------------- preprocess */

typedef struct {
    uint64_t deg;
    uint32_t id;
} degid;

static int cmp_degid(const void *x, const void *y) {
    const degid *a = (const degid *)x, *b = (const degid *)y;
    if (a->deg != b->deg) return a->deg < b->deg ? -1 : 1;
    if (a->id != b->id) return a->id < b->id ? -1 : 1;
    return 0;
}

static int cmp_u32(const void *x, const void *y) {
    uint32_t a = *(const uint32_t *)x, b = *(const uint32_t *)y;
    return a < b ? -1 : (a > b);
}

/* preprocess, tc.cu:687-787. */
int orc_preprocess(const orc_graph *in, orc_graph *out) {
    memset(out, 0, sizeof(*out));
    uint32_t n = in->n;
    degid *vs = (degid *)malloc((n ? n : 1) * sizeof(degid));
    uint32_t *rank = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    if (!vs || !rank) { free(vs); free(rank); return ORC_ERR_NOMEM; }
    /* tc.cu:691-716: records (id, degree), sorted ascending by degree; the
     * (degree,id) key makes the order that of a stable sort. */
    for (uint32_t v = 0; v < n; v++) { vs[v].id = v; vs[v].deg = in->rowPtr[v + 1] - in->rowPtr[v]; }
    qsort(vs, n, sizeof(degid), cmp_degid);
    /* tc.cu:718-723: reverse[old id] = new rank. */
    for (uint32_t r = 0; r < n; r++) rank[vs[r].id] = r;
    out->n = n;
    out->m = in->m / 2; /* tc.cu:729 */
    out->rowPtr = (uint64_t *)malloc(((uint64_t)n + 1) * sizeof(uint64_t));
    out->colInd = (uint32_t *)malloc((out->m ? out->m : 1) * sizeof(uint32_t));
    if (!out->rowPtr || !out->colInd) { free(vs); free(rank); orc_free(out); return ORC_ERR_NOMEM; }
    /* tc.cu:738-758: orient low rank -> high rank, sort each row. */
    uint64_t k = 0;
    out->rowPtr[0] = 0;
    for (uint32_t r = 0; r < n; r++) {
        uint32_t id = vs[r].id;
        uint64_t start = k;
        for (uint64_t j = in->rowPtr[id]; j < in->rowPtr[id + 1]; j++) {
            uint32_t wn = rank[in->colInd[j]];
            if (wn > r) out->colInd[k++] = wn;
        }
        qsort(out->colInd + start, k - start, sizeof(uint32_t), cmp_u32);
        out->rowPtr[r + 1] = k;
    }
    free(vs);
    free(rank);
    if (k != out->m) { orc_free(out); return ORC_ERR_FORMAT; } /* input was not symmetric */
    return ORC_OK;
}

/* --------------------------------------------------------- wedge prefix etc */

uint64_t orc_wedge_prefix(const orc_graph *g, uint64_t *ws) {
    ws[0] = 0;
    for (uint32_t v = 0; v < g->n; v++) {
        uint64_t d = g->rowPtr[v + 1] - g->rowPtr[v];
        ws[v + 1] = ws[v] + (d < 2 ? 0 : binom2(d));
    }
    return ws[g->n];
}

/* build_binary_search_cache, tc.cu:306-313. */
static void tree_rec(const uint64_t *src, uint64_t *tree, uint32_t level, uint32_t i, uint32_t s, uint32_t e) {
    if (level >= ORC_TREE_LEVELS) return;
    uint32_t mid = (uint32_t)(((uint64_t)s + e) / 2);
    tree[i] = src[mid];
    tree_rec(src, tree, level + 1, 2 * i + 1, s, mid);
    tree_rec(src, tree, level + 1, 2 * i + 2, mid + 1, e);
}

void orc_search_tree(const uint64_t *ws, uint32_t n, uint64_t *tree) { tree_rec(ws, tree, 0, 0, 0, n); }

/* tc.cu:117-137. */
uint32_t orc_search_closest(const uint64_t *list, uint32_t start, uint32_t end, uint64_t target) {
    uint32_t s = start, e = end;
    while (s < e) {
        uint32_t mid = (uint32_t)(((uint64_t)s + e) >> 1);
        if (list[mid] < target + 1) s = mid + 1;
        else e = mid;
    }
    uint32_t r = s > 0 ? s - 1 : 0;
    return r > start ? r : start;
}

/* tc.cu:139-170. */
uint32_t orc_search_closest_tree(const uint64_t *tree, const uint64_t *list, uint32_t start, uint32_t end,
                                 uint64_t target) {
    uint32_t gs = start, ge = end, ci = 0;
    for (uint32_t it = 0; it < ORC_TREE_LEVELS; it++) {
        uint64_t mid = tree[ci];
        uint32_t gmid = (uint32_t)(((uint64_t)gs + ge) >> 1);
        ci = 2 * ci + 1;
        if (mid < target + 1) { ci += 1; gs = gmid + 1; }
        else ge = gmid;
    }
    uint32_t r = gs > 0 ? gs - 1 : 0;
    gs = r > start ? r : start;
    return orc_search_closest(list, gs, ge, target);
}

/* tc.cu:209-212, integer-exact: with k = C(d,2) - i_v wedges left (k >= 1),
 * the row is w_i = d-1-q for the largest q with C(q,2) < k (binary search on
 * q, no floating point), and u_i follows tc.cu:212. */
void orc_decode(uint64_t d, uint64_t i_v, uint64_t *w_i, uint64_t *u_i) {
    uint64_t C = binom2(d);
    uint64_t k = C - i_v;
    uint64_t lo = 1, hi = d - 1; /* C(1,2)=0 < k always; C(d,2) >= k */
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo + 1) / 2;
        if (binom2(mid) < k) lo = mid;
        else hi = mid - 1;
    }
    uint64_t w = d - 1 - lo;
    *w_i = w;
    *u_i = i_v + w + 1 + (binom2(d - w) - C);
}

/* tc.cu:211-212 as written (64-bit variant experiments/tc_big_graphs.cu:211-212). */
void orc_decode_ref_fp(uint64_t d, uint64_t i_v, uint64_t *w_i, uint64_t *u_i) {
    uint64_t wedges = binom2(d);
    uint32_t w = (uint32_t)(d - 2 - (uint32_t)(sqrt(((double)(wedges - i_v) - 0.875) * 2) - 0.5));
    *w_i = w;
    *u_i = i_v + w + 1 + (binom2(d - w) - wedges);
}

/* Exhaustive decode check for d in [d_lo, d_hi]: enumerates the pairs of the
 * strict upper triangle row by row (the order tc.cu:215-240 steps through)
 * and compares orc_decode and, when fp != 0, the reference formula.  Returns
 * 0 if all agree, else the first bad d (with its index in *bad_i). */
uint64_t orc_decode_check(uint64_t d_lo, uint64_t d_hi, int fp, uint64_t *bad_i) {
    for (uint64_t d = d_lo < 2 ? 2 : d_lo; d <= d_hi; d++) {
        uint64_t i = 0;
        for (uint64_t w = 0; w + 1 < d; w++)
            for (uint64_t u = w + 1; u < d; u++, i++) {
                uint64_t a, b;
                orc_decode(d, i, &a, &b);
                if (a != w || b != u) { *bad_i = i; return d; }
                if (fp) {
                    orc_decode_ref_fp(d, i, &a, &b);
                    if (a != w || b != u) { *bad_i = i; return d; }
                }
            }
    }
    return 0;
}

/* ------------------------------------------------------------------ bitmap */

uint64_t orc_bitmap_bits(uint32_t a) { return binom2(a); }
/* tc.cu:331 allocates bits/32 words (floor); the index range needs the
 * ceiling when a(a-1)/2 is not a multiple of 32 (e.g. a = 32). */
uint64_t orc_bitmap_words(uint32_t a) { return (binom2(a) + 31u) / 32u; }

/* tc.cu:263 / tc.cu:338. */
uint64_t orc_bitmap_index(uint32_t n, uint32_t a, uint32_t w, uint32_t u) {
    return (binom2(a) - binom2((uint64_t)n - w)) + u - w - 1;
}

/* tc.cu:329-346. */
void orc_bitmap_build(const orc_graph *g, uint32_t a, uint32_t *words) {
    uint32_t n = g->n;
    uint32_t top = n < a ? n : a;
    for (uint32_t v = n - top; v < n; v++)
        for (uint64_t i = g->rowPtr[v]; i < g->rowPtr[v + 1]; i++) {
            uint64_t idx = orc_bitmap_index(n, a, v, g->colInd[i]);
            words[idx >> 5] |= 1u << (idx & 31u);
        }
}

/* ------------------------------------------------------- wedge-parallel TC */

/* linear_search_GPU tc.cu:89-99 / binary_search_GPU tc.cu:101-114. */
static int list_contains(const uint32_t *list, uint64_t start, uint64_t end, uint32_t target) {
    if (end - start < 2) {
        for (uint64_t i = start; i < end; i++) {
            if (list[i] == target) return 1;
            if (list[i] > target) break;
        }
        return 0;
    }
    uint64_t s = start, e = end;
    while (s < e) {
        uint64_t mid = (s + e) >> 1;
        if (list[mid] == target) return 1;
        if (list[mid] < target) s = mid + 1;
        else e = mid;
    }
    return 0;
}

int orc_tc_wedge(const orc_graph *g, uint32_t spread, uint32_t a, uint64_t w_begin, uint64_t w_end,
                 uint64_t *count) {
    if (spread == 0) return ORC_ERR_ARG;
    const uint32_t B = 128; /* tc.cu:394 */
    uint32_t n = g->n;
    uint64_t *ws = (uint64_t *)malloc(((uint64_t)n + 1) * sizeof(uint64_t));
    uint64_t tree[ORC_TREE_SIZE];
    uint64_t bm_words = orc_bitmap_words(a);
    uint32_t *bm = (uint32_t *)calloc(bm_words ? bm_words : 1, sizeof(uint32_t));
    uint32_t *ssrc = (uint32_t *)malloc((uint64_t)B * spread * sizeof(uint32_t));
    uint32_t *sdst = (uint32_t *)malloc((uint64_t)B * spread * sizeof(uint32_t));
    if (!ws || !bm || !ssrc || !sdst) { free(ws); free(bm); free(ssrc); free(sdst); return ORC_ERR_NOMEM; }
    uint64_t W = orc_wedge_prefix(g, ws);
    if (w_end > W) w_end = W;
    orc_search_tree(ws, n, tree);
    orc_bitmap_build(g, a, bm);
    uint64_t bm_size = binom2(a);
    uint32_t thr = (n > a ? n : a) - a;
    const uint64_t *Ap = g->rowPtr;
    const uint32_t *Ai = g->colInd;
    uint64_t total = 0;
    for (uint64_t blk_base = w_begin; blk_base < w_end; blk_base += (uint64_t)B * spread) {
        /* phase 1, tc.cu:192-244 */
        for (uint32_t t = 0; t < B; t++) {
            uint64_t i_start = blk_base + (uint64_t)t * spread;
            uint64_t stop = i_start + spread < w_end ? i_start + spread : w_end;
            uint32_t v = 0, w = 0;
            uint64_t vb = 0, ve = 0, d = 0, wedges = 0, i_v = 0, w_i = 0, u_i = 0;
            uint32_t s_i = t * spread;
            for (uint64_t i = i_start; i < stop; i++, i_v++, s_i++) {
                if (i == i_start) {
                    v = orc_search_closest_tree(tree, ws, 0, n, i_start);
                    vb = Ap[v];
                    ve = Ap[v + 1];
                    d = ve - vb;
                    wedges = binom2(d);
                    i_v = i_start - ws[v];
                    orc_decode(d, i_v, &w_i, &u_i);
                    w = Ai[vb + w_i];
                } else if (i_v >= wedges) {
                    do {
                        v++;
                        vb = ve;
                        ve = Ap[v + 1];
                        d = ve - vb;
                    } while (d < 2);
                    wedges = binom2(d);
                    i_v = 0;
                    w_i = 0;
                    u_i = 1;
                    w = Ai[vb];
                } else {
                    u_i++;
                    if (u_i >= d) {
                        w_i++;
                        u_i = w_i + 1;
                        w = Ai[vb + w_i];
                    }
                }
                ssrc[s_i] = w;
                sdst[s_i] = Ai[vb + u_i];
            }
        }
        /* phase 2, tc.cu:251-287 */
        uint64_t limit = w_end - blk_base;
        for (uint64_t s_i = 0; s_i < (uint64_t)B * spread && s_i < limit; s_i++) {
            uint32_t w = ssrc[s_i], u = sdst[s_i];
            if (w >= thr) {
                uint64_t idx = (bm_size - binom2((uint64_t)n - w)) + u - w - 1;
                total += (bm[idx >> 5] >> (idx & 31u)) & 1u;
            } else {
                total += (uint64_t)list_contains(Ai, Ap[w], Ap[w + 1], u);
            }
        }
    }
    free(ws);
    free(bm);
    free(ssrc);
    free(sdst);
    *count = total;
    return ORC_OK;
}

/* ------------------------------------------------------- arrow-wedge TC */

/* Arrow-wedge prefix, experiments/tc_arrow.cu:345-368: vi[0] = 0 and, for
 * each edge i of row v, vi[i+1] = vi[i] + d+(colInd[i]) when d+(v) >= 2,
 * else vi[i+1] = vi[i].  Returns the arrow-wedge total vi[m]. */
uint64_t orc_arrow_prefix(const orc_graph *g, uint64_t *vi) {
    vi[0] = 0;
    for (uint32_t v = 0; v < g->n; v++) {
        uint64_t vb = g->rowPtr[v], ve = g->rowPtr[v + 1];
        for (uint64_t i = vb; i < ve; i++) {
            uint32_t w = g->colInd[i];
            uint64_t dw = ve - vb >= 2 ? g->rowPtr[w + 1] - g->rowPtr[w] : 0;
            vi[i + 1] = vi[i] + dw;
        }
    }
    return vi[g->m];
}

/* tc_arrow_GPU_kernel (experiments/tc_arrow.cu:173-289) executed
 * sequentially: arrow wedge i of edge e = (v -> w) names u =
 * colInd[rowPtr[w] + i - vi[e]], and counts when u is in adj+(v): a bitmap
 * test for v in the top-a ranks (tc_arrow.cu:256-266, the same index as
 * tc.cu:263 with v in place of w), else a list search (tc_arrow.cu:271-284).
 * Every wedge is visited once, in index order; the block/lane split of the
 * kernel only changes which thread visits it. */
int orc_tc_arrow(const orc_graph *g, uint32_t a, uint64_t *arrow_wedges, uint64_t *count) {
    const uint32_t n = g->n;
    uint64_t *vi = (uint64_t *)malloc((g->m + 1) * sizeof(uint64_t));
    uint64_t bm_words = orc_bitmap_words(a);
    uint32_t *bm = (uint32_t *)calloc(bm_words ? bm_words : 1, sizeof(uint32_t));
    if (!vi || !bm) { free(vi); free(bm); return ORC_ERR_NOMEM; }
    uint64_t A = orc_arrow_prefix(g, vi);
    if (a) orc_bitmap_build(g, a, bm);
    const uint32_t thr = a ? (n > a ? n - a : 0) : n;
    const uint64_t bm_size = orc_bitmap_bits(a);
    const uint64_t *Ap = g->rowPtr;
    const uint32_t *Ai = g->colInd;
    uint64_t total = 0;
    for (uint32_t v = 0; v < n; v++) {
        for (uint64_t e = Ap[v]; e < Ap[v + 1]; e++) {
            uint32_t w = Ai[e];
            for (uint64_t k = 0; k < vi[e + 1] - vi[e]; k++) {
                uint32_t u = Ai[Ap[w] + k];
                if (v >= thr) {
                    uint64_t idx = (bm_size - binom2((uint64_t)n - v)) + u - v - 1;
                    total += (bm[idx >> 5] >> (idx & 31u)) & 1u;
                } else {
                    total += (uint64_t)list_contains(Ai, Ap[v], Ap[v + 1], u);
                }
            }
        }
    }
    free(vi);
    free(bm);
    if (arrow_wedges) *arrow_wedges = A;
    *count = total;
    return ORC_OK;
}

/* ------------------------------------------- vertex/edge-parallel styles */

static uint32_t probe_has(const orc_graph *g, const uint32_t *bm, uint32_t thr, uint64_t bm_size, uint32_t x,
                          uint64_t lo, uint64_t hi, uint32_t u) {
    if (x >= thr) {
        uint64_t idx = (bm_size - binom2((uint64_t)g->n - x)) + u - x - 1;
        return (bm[idx >> 5] >> (idx & 31u)) & 1u;
    }
    return (uint32_t)list_contains(g->colInd, lo, hi, u);
}

/* The per-edge work of tc_vertex_{outgoing,arrow,mixed}_GPU_kernel
 * (experiments/tc_vertex.cu:113-274) and tc_edge_*_GPU_kernel
 * (experiments/tc_edge.cu:179-352), done for every edge in order: the vertex-
 * and edge-parallel kernels only differ in which thread does an edge.
 * style 0 outgoing, 1 arrow, 2 mixed (tc_vertex.cu:57). */
int orc_tc_style(const orc_graph *g, int style, uint32_t a, uint64_t *count) {
    if (style < 0 || style > 2) return ORC_ERR_ARG;
    const uint32_t n = g->n;
    uint64_t bm_words = orc_bitmap_words(a);
    uint32_t *bm = (uint32_t *)calloc(bm_words ? bm_words : 1, sizeof(uint32_t));
    if (!bm) return ORC_ERR_NOMEM;
    if (a) orc_bitmap_build(g, a, bm);
    const uint32_t thr = a ? (n > a ? n - a : 0) : n;
    const uint64_t bm_size = orc_bitmap_bits(a);
    const uint64_t *Ap = g->rowPtr;
    const uint32_t *Ai = g->colInd;
    uint64_t total = 0;
    for (uint32_t v = 0; v < n; v++) {
        uint64_t vb = Ap[v], ve = Ap[v + 1];
        for (uint64_t i = vb; i < ve; i++) {
            uint32_t w = Ai[i];
            uint64_t wb = Ap[w], we = Ap[w + 1];
            int walk_v = style == 0 || (style == 2 && ve - (i + 1) <= we - wb);
            if (walk_v) {
                for (uint64_t j = i + 1; j < ve; j++) total += probe_has(g, bm, thr, bm_size, w, wb, we, Ai[j]);
            } else {
                /* mixed searches after w (tc_vertex.cu:258; tc_edge.cu:337 searches the whole row,
                 * the same membership since u > w), arrow the whole row (tc_vertex.cu:187) */
                uint64_t lo = style == 2 ? i + 1 : vb;
                for (uint64_t j = wb; j < we; j++) total += probe_has(g, bm, thr, bm_size, v, lo, ve, Ai[j]);
            }
        }
    }
    free(bm);
    *count = total;
    return ORC_OK;
}

/* ------------------------------------------------------- merge CPU baseline */

typedef struct {
    const orc_graph *g;
    atomic_uint_fast64_t *next;
    uint64_t v_end;
    uint64_t sum;
} merge_arg;

static void *merge_worker(void *p) {
    merge_arg *a = (merge_arg *)p;
    const orc_graph *g = a->g;
    const uint64_t *Ap = g->rowPtr;
    const uint32_t *Ai = g->colInd;
    const uint64_t chunk = 256;
    uint64_t sum = 0;
    for (;;) {
        uint64_t v0 = atomic_fetch_add(a->next, chunk);
        if (v0 >= a->v_end) break;
        uint64_t v1 = v0 + chunk < a->v_end ? v0 + chunk : a->v_end;
        for (uint64_t v = v0; v < v1; v++) {
            uint64_t vb = Ap[v], ve = Ap[v + 1];
            for (uint64_t j = vb; j < ve; j++) {
                uint32_t w = Ai[j];
                uint64_t x = j + 1, y = Ap[w], ye = Ap[w + 1];
                while (x < ve && y < ye) {
                    uint32_t p = Ai[x], q = Ai[y];
                    if (p == q) { sum++; x++; y++; }
                    else if (p < q) x++;
                    else y++;
                }
            }
        }
    }
    a->sum = sum;
    return NULL;
}

uint64_t orc_tc_merge_range(const orc_graph *g, uint32_t v0, uint32_t v1, int threads) {
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads < 1) threads = 1;
    if (v1 > g->n) v1 = g->n;
    atomic_uint_fast64_t next;
    atomic_init(&next, v0);
    pthread_t *th = (pthread_t *)malloc((size_t)threads * sizeof(pthread_t));
    merge_arg *args = (merge_arg *)calloc((size_t)threads, sizeof(merge_arg));
    for (int t = 0; t < threads; t++) {
        args[t].g = g;
        args[t].next = &next;
        args[t].v_end = v1;
        pthread_create(&th[t], NULL, merge_worker, &args[t]);
    }
    uint64_t total = 0;
    for (int t = 0; t < threads; t++) {
        pthread_join(th[t], NULL);
        total += args[t].sum;
    }
    free(th);
    free(args);
    return total;
}

uint64_t orc_tc_merge(const orc_graph *g, int threads) { return orc_tc_merge_range(g, 0, g->n, threads); }

/* ---------------------------------------------------------- alg bytes */

static uint64_t bitlen(uint64_t x) { return x ? 64u - (uint64_t)__builtin_clzll(x) : 0u; }

void orc_alg_bytes(const orc_graph *g, uint32_t o, double *b_alg, double *b_min) {
    double base = (double)o * ((double)g->n + 1) + 4.0 * (double)g->m + 8.0 * ((double)g->n + 1);
    double probes = 0, W = 0;
    for (uint32_t v = 0; v < g->n; v++) {
        uint64_t vb = g->rowPtr[v], d = g->rowPtr[v + 1] - vb;
        for (uint64_t j = 0; j < d; j++) {
            uint32_t w = g->colInd[vb + j];
            uint64_t cnt = d - 1 - j; /* wedges of v whose first leg is w */
            uint64_t dw = g->rowPtr[w + 1] - g->rowPtr[w];
            probes += (double)cnt * (4.0 + 4.0 * (double)bitlen(dw));
            W += (double)cnt;
        }
    }
    *b_alg = base + probes;
    *b_min = base + 8.0 * W;
}
