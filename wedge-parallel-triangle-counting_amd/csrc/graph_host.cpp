// graph_host.cpp -- host side of the drop-in: graph reader, symmetric CSR
// builder, degree-order preprocessing and seeded graph generators.
//
// Semantics follow the reference's read_graph (tc.cu:561-685) and preprocess
// (tc.cu:687-787); the construction is this library's own: a two-pass
// counting scatter into rows, per-row sort/dedup and a counting sort by
// degree, all multi-threaded and deterministic for any thread count.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <unordered_set>

#include "host_util.h"
#include "wtc.h"

namespace wtc {

int host_threads() {
    static const int nt = [] {
        const char *e = std::getenv("WTC_HOST_THREADS");
        if (!e || !*e) e = std::getenv("OMP_NUM_THREADS");
        int v = e ? std::atoi(e) : 0;
        if (v <= 0) v = (int)std::thread::hardware_concurrency();
        if (v <= 0) v = 1;
        return std::min(v, 64);
    }();
    return nt;
}

int validate_oriented_rows(uint64_t n, uint64_t m, const uint64_t *rowPtr, const uint32_t *colInd, uint64_t v0,
                           uint64_t v1) {
    return validate_oriented_rows_t(n, m, rowPtr, colInd, v0, v1);
}

// Runs f(t) on T threads (t = 0..T-1) and joins.
template <class F>
static void run_threads(int T, F &&f) {
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) th.emplace_back(f, t);
    f(0);
    for (auto &x : th) x.join();
}

// Builds the symmetric, deduplicated CSR of the undirected pairs produced by
// get(i, &s, &d) for i < np (get returns false for pairs to skip).  ids must
// be < n_cap; the result has n = (largest id with an edge) + 1, or 1 when
// there is no edge (the reference's max_vertex starts at 0, tc.cu:587,625).
//
// Both directions of every pair become (src<<32|dst) keys, scattered by the
// top 16 bits of src into buckets (per-thread histograms over a static split
// of the pairs, so no atomics), then each bucket is sorted and deduplicated
// on its own; buckets cover ascending src ranges, so their concatenation is
// the CSR.
template <class Get>
static int build_symmetric(uint64_t n_cap, uint64_t np, Get get, wtc_big_graph **out) {
    if (n_cap == 0) n_cap = 1;
    if (n_cap > 0xffffffffULL) return WTC_ERR_FORMAT;
    int bits = 1;
    while ((1ull << bits) < n_cap) bits++;
    const int kb = std::min(bits, 16), shift = bits - kb;
    const uint64_t NB = 1ull << kb;
    const int T = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)host_threads(), (np >> 16) + 1));
    const uint64_t per = (np + T - 1) / T;
    std::vector<uint64_t> hist((size_t)T * NB, 0), bstart(NB + 1), ulen(NB);
    run_threads(T, [&](int t) {
        uint64_t *h = hist.data() + (size_t)t * NB;
        uint32_t s, d;
        for (uint64_t i = (uint64_t)t * per, e = std::min(np, (uint64_t)(t + 1) * per); i < e; i++)
            if (get(i, s, d)) {
                h[s >> shift]++;
                h[d >> shift]++;
            }
    });
    uint64_t acc = 0;
    for (uint64_t b = 0; b < NB; b++) {
        bstart[b] = acc;
        for (int t = 0; t < T; t++) {
            uint64_t v = hist[(size_t)t * NB + b];
            hist[(size_t)t * NB + b] = acc;
            acc += v;
        }
    }
    bstart[NB] = acc;
    uint64_t *keys = alloc_n<uint64_t>(acc);
    uint32_t *deg = static_cast<uint32_t *>(std::calloc(n_cap, sizeof(uint32_t)));
    if (!keys || !deg) {
        std::free(keys);
        std::free(deg);
        return WTC_ERR_NOMEM;
    }
    run_threads(T, [&](int t) {
        uint64_t *h = hist.data() + (size_t)t * NB;
        uint32_t s, d;
        for (uint64_t i = (uint64_t)t * per, e = std::min(np, (uint64_t)(t + 1) * per); i < e; i++)
            if (get(i, s, d)) {
                keys[h[s >> shift]++] = ((uint64_t)s << 32) | d;
                keys[h[d >> shift]++] = ((uint64_t)d << 32) | s;
            }
    });
    par_chunks(NB, 1, [&](uint64_t b0, uint64_t b1, int) {
        for (uint64_t b = b0; b < b1; b++) {
            uint64_t *k0 = keys + bstart[b], *k1 = keys + bstart[b + 1];
            std::sort(k0, k1);
            k1 = std::unique(k0, k1);
            ulen[b] = (uint64_t)(k1 - k0);
            for (uint64_t *k = k0; k < k1; k++) deg[*k >> 32]++;  // the bucket owns these rows
        }
    });
    uint64_t n = 1;
    for (uint64_t v = n_cap; v-- > 0;)
        if (deg[v]) {
            n = v + 1;
            break;
        }
    wtc_big_graph *g = static_cast<wtc_big_graph *>(std::calloc(1, sizeof(wtc_big_graph)));
    uint64_t *rp = alloc_n<uint64_t>(n + 1);
    uint32_t *ci = nullptr;
    if (g && rp) {
        rp[0] = 0;
        for (uint64_t v = 0; v < n; v++) rp[v + 1] = rp[v] + deg[v];
        ci = alloc_n<uint32_t>(rp[n]);
    }
    if (!g || !rp || !ci) {
        std::free(g);
        std::free(rp);
        std::free(ci);
        std::free(keys);
        std::free(deg);
        return WTC_ERR_NOMEM;
    }
    par_chunks(NB, 16, [&](uint64_t b0, uint64_t b1, int) {
        for (uint64_t b = b0; b < b1; b++) {
            if (ulen[b] == 0) continue;
            const uint64_t *k = keys + bstart[b];
            uint32_t *dst = ci + rp[k[0] >> 32];
            for (uint64_t j = 0; j < ulen[b]; j++) dst[j] = (uint32_t)k[j];
        }
    });
    std::free(keys);
    std::free(deg);
    g->numVertices = (uint32_t)n;
    g->numEdges = rp[n];
    g->rowPtr = rp;
    g->colInd = ci;
    *out = g;
    return WTC_OK;
}

int build_symmetric_tokens(uint64_t n_cap, const uint32_t *tok, uint64_t np, uint32_t shift, wtc_big_graph **out) {
    return build_symmetric(n_cap, np, [tok, shift](uint64_t i, uint32_t &s, uint32_t &d) {
        s = tok[2 * i] - shift;
        d = tok[2 * i + 1] - shift;
        return s != d;
    }, out);
}

}  // namespace wtc

using namespace wtc;

extern "C" {

const char *wtc_strerror(int s) {
    switch (s) {
        case WTC_OK: return "ok";
        case WTC_ERR_ARG: return "invalid argument";
        case WTC_ERR_IO: return "unable to open graph file";
        case WTC_ERR_FORMAT: return "graph file has no edge line or ids out of range";
        case WTC_ERR_ZERO_ID: return "zero vertex id detected but -z was not set";
        case WTC_ERR_NOMEM: return "failed to allocate host memory";
        case WTC_ERR_HIP: return "HIP runtime error";
        case WTC_ERR_RANGE: return "graph too large for the 32-bit path";
        case WTC_ERR_NODEV: return "no usable HIP device";
        case WTC_ERR_ADMIT: return "request exceeds the device memory budget";
        case WTC_ERR_CSR: return "CSR failed validation";
        case WTC_ERR_BUSY: return "server queue full or stopped";
        default: return "unknown error";
    }
}

int wtc_graph_from_pairs(const uint32_t *pairs, uint64_t npairs, wtc_big_graph **out) {
    if (!out || (npairs && !pairs)) return WTC_ERR_ARG;
    *out = nullptr;
    uint32_t mx = 0;
    for (uint64_t i = 0; i < npairs; i++)
        if (pairs[2 * i] != pairs[2 * i + 1]) mx = std::max(mx, std::max(pairs[2 * i], pairs[2 * i + 1]));
    if ((uint64_t)mx + 1 > 0xffffffffULL) return WTC_ERR_FORMAT;
    return build_symmetric((uint64_t)mx + 1, npairs,
                           [pairs](uint64_t i, uint32_t &s, uint32_t &d) {
                               s = pairs[2 * i];
                               d = pairs[2 * i + 1];
                               return s != d;
                           },
                           out);
}

void wtc_free_graph(wtc_big_graph *g) {
    if (!g) return;
    std::free(g->rowPtr);
    std::free(g->colInd);
    std::free(g);
}

void wtc_free_graph32(wtc_graph *g) {
    if (!g) return;
    std::free(g->rowPtr);
    std::free(g->colInd);
    std::free(g);
}

void wtc_free_pairs(uint32_t *pairs) { std::free(pairs); }

// preprocess, tc.cu:687-787.
int wtc_preprocess(const wtc_big_graph *in, int preprocess_style, wtc_big_graph **out) {
    if (!in || !out) return WTC_ERR_ARG;
    if (preprocess_style < WTC_PREPROCESS_CPU || preprocess_style > WTC_PREPROCESS_GPU_CONSTRAINED)
        return WTC_ERR_ARG;
    *out = nullptr;
    // -p 1/2: degree sort, orientation and row sort on the GPU (tc.cu:697-716, 763-784)
    if (preprocess_style != WTC_PREPROCESS_CPU) return wtc_gpu_orient_default(in, out);
    const uint64_t n = in->numVertices;
    const uint64_t *rp = in->rowPtr;
    // Stable ascending order by degree: counting sort, scattered in id order.
    uint64_t maxdeg = 0;
    for (uint64_t v = 0; v < n; v++) maxdeg = std::max(maxdeg, rp[v + 1] - rp[v]);
    uint64_t *bucket = static_cast<uint64_t *>(std::calloc(maxdeg + 2, sizeof(uint64_t)));
    uint32_t *order = alloc_n<uint32_t>(n);  // rank -> old id
    uint32_t *rank = alloc_n<uint32_t>(n);   // old id -> rank
    uint64_t *dplus = alloc_n<uint64_t>(n + 1);
    auto fail = [&](int rc) {
        std::free(bucket);
        std::free(order);
        std::free(rank);
        std::free(dplus);
        return rc;
    };
    if (!bucket || !order || !rank || !dplus) return fail(WTC_ERR_NOMEM);
    for (uint64_t v = 0; v < n; v++) bucket[rp[v + 1] - rp[v] + 1]++;
    for (uint64_t k = 1; k <= maxdeg + 1; k++) bucket[k] += bucket[k - 1];
    for (uint64_t v = 0; v < n; v++) {
        uint64_t r = bucket[rp[v + 1] - rp[v]]++;
        order[r] = (uint32_t)v;
        rank[v] = (uint32_t)r;
    }
    // out-degree per rank: neighbours of higher rank
    const uint32_t *ci = in->colInd;
    par_chunks(n, 4096, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t r = b; r < e; r++) {
            uint32_t id = order[r];
            uint64_t c = 0;
            for (uint64_t j = rp[id]; j < rp[id + 1]; j++) c += rank[ci[j]] > r;
            dplus[r] = c;
        }
    });
    uint64_t *orp = alloc_n<uint64_t>(n + 1);
    if (!orp) return fail(WTC_ERR_NOMEM);
    orp[0] = 0;
    for (uint64_t r = 0; r < n; r++) orp[r + 1] = orp[r] + dplus[r];
    if (orp[n] != in->numEdges / 2) {  // input not symmetric
        std::free(orp);
        return fail(WTC_ERR_CSR);
    }
    uint32_t *oci = alloc_n<uint32_t>(orp[n]);
    wtc_big_graph *g = static_cast<wtc_big_graph *>(std::calloc(1, sizeof(wtc_big_graph)));
    if (!oci || !g) {
        std::free(orp);
        std::free(oci);
        std::free(g);
        return fail(WTC_ERR_NOMEM);
    }
    par_chunks(n, 4096, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t r = b; r < e; r++) {
            uint32_t id = order[r];
            uint32_t *dst = oci + orp[r];
            uint64_t k = 0;
            for (uint64_t j = rp[id]; j < rp[id + 1]; j++) {
                uint32_t wn = rank[ci[j]];
                if (wn > r) dst[k++] = wn;
            }
            std::sort(dst, dst + k);
        }
    });
    fail(WTC_OK);
    g->numVertices = (uint32_t)n;
    g->numEdges = orp[n];
    g->rowPtr = orp;
    g->colInd = oci;
    *out = g;
    return WTC_OK;
}

int wtc_narrow_graph(const wtc_big_graph *g, wtc_graph **out) {
    if (!g || !out) return WTC_ERR_ARG;
    *out = nullptr;
    if (g->numEdges > 0xffffffffULL) return WTC_ERR_RANGE;
    wtc_graph *s = static_cast<wtc_graph *>(std::calloc(1, sizeof(wtc_graph)));
    uint32_t *rp = alloc_n<uint32_t>((uint64_t)g->numVertices + 1);
    uint32_t *ci = alloc_n<uint32_t>(g->numEdges);
    if (!s || !rp || !ci) {
        std::free(s);
        std::free(rp);
        std::free(ci);
        return WTC_ERR_NOMEM;
    }
    for (uint64_t v = 0; v <= g->numVertices; v++) rp[v] = (uint32_t)g->rowPtr[v];
    std::memcpy(ci, g->colInd, g->numEdges * sizeof(uint32_t));
    s->numVertices = g->numVertices;
    s->numEdges = (uint32_t)g->numEdges;
    s->rowPtr = rp;
    s->colInd = ci;
    *out = s;
    return WTC_OK;
}

int wtc_validate_oriented(const wtc_big_graph *g) {
    if (!g || !g->rowPtr || (g->numEdges && !g->colInd)) return WTC_ERR_ARG;
    return validate_oriented_par(g->numVertices, g->numEdges, g->rowPtr, g->colInd);
}

uint64_t wtc_wedge_total(const wtc_big_graph *g) {
    uint64_t W = 0;
    for (uint64_t v = 0; v < g->numVertices; v++) W += binom2(g->rowPtr[v + 1] - g->rowPtr[v]);
    return W;
}

void wtc_alg_bytes(const wtc_big_graph *g, uint32_t o, double *b_alg, double *b_min) {
    const uint64_t n = g->numVertices;
    const double base = (double)o * (double)(n + 1) + 4.0 * (double)g->numEdges + 8.0 * (double)(n + 1);
    const uint64_t chunk = 1 << 14, nch = (n + chunk - 1) / chunk;
    std::vector<double> pw(nch, 0.0), ww(nch, 0.0);
    par_chunks(n, chunk, [&](uint64_t b, uint64_t e, int) {
        uint64_t probe_words = 0, wedges = 0;  // exact integer sums per chunk
        for (uint64_t v = b; v < e; v++) {
            const uint64_t vb = g->rowPtr[v], d = g->rowPtr[v + 1] - vb;
            for (uint64_t j = 0; j < d; j++) {
                const uint32_t w = g->colInd[vb + j];
                const uint64_t dw = g->rowPtr[w + 1] - g->rowPtr[w];
                const uint64_t lv = dw ? 64u - (uint64_t)__builtin_clzll(dw) : 0u;  // ceil(log2(dw+1))
                const uint64_t cnt = d - 1 - j;
                probe_words += cnt * (1 + lv);
                wedges += cnt;
            }
        }
        pw[b / chunk] = 4.0 * (double)probe_words;
        ww[b / chunk] = (double)wedges;
    });
    double probes = 0, W = 0;
    for (uint64_t c = 0; c < nch; c++) {
        probes += pw[c];
        W += ww[c];
    }
    if (b_alg) *b_alg = base + probes;
    if (b_min) *b_min = base + 8.0 * W;
}

// ---------------------------------------------------------------- generators

struct RmatGen {
    uint32_t scale;
    uint64_t seed;
    uint64_t ta, tab, tabc;  // cumulative quadrant thresholds on 32-bit draws
    void operator()(uint64_t i, uint32_t &s, uint32_t &d) const {
        uint32_t src = 0, dst = 0;
        uint64_t r = 0;
        for (uint32_t l = 0; l < scale; l++) {
            if ((l & 1u) == 0) r = rng(seed, 0x524d4154ULL, i * 64 + l);
            uint64_t x = (l & 1u) ? (r >> 32) : (r & 0xffffffffULL);
            uint32_t sb = x >= tab, db = (x >= ta && x < tab) || x >= tabc;
            src = (src << 1) | sb;
            dst = (dst << 1) | db;
        }
        s = src;
        d = dst;
    }
};

static int rmat_setup(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed, RmatGen &g,
                      uint64_t &np) {
    if (scale == 0 || scale > 31 || ef == 0) return WTC_ERR_ARG;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return WTC_ERR_ARG;
    const double two32 = 4294967296.0;
    g.scale = scale;
    g.seed = seed;
    g.ta = (uint64_t)(a * two32);
    g.tab = (uint64_t)((a + b) * two32);
    g.tabc = (uint64_t)((a + b + c) * two32);
    np = (uint64_t)ef << scale;
    return WTC_OK;
}

int wtc_gen_rmat(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed, wtc_big_graph **out) {
    if (!out) return WTC_ERR_ARG;
    *out = nullptr;
    RmatGen gen;
    uint64_t np;
    int rc = rmat_setup(scale, ef, a, b, c, seed, gen, np);
    if (rc) return rc;
    return build_symmetric(1ULL << scale, np,
                           [&gen](uint64_t i, uint32_t &s, uint32_t &d) {
                               gen(i, s, d);
                               return s != d;
                           },
                           out);
}

int wtc_gen_rmat_pairs(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed, uint32_t **pairs,
                       uint64_t *npairs) {
    if (!pairs || !npairs) return WTC_ERR_ARG;
    RmatGen gen;
    uint64_t np;
    int rc = rmat_setup(scale, ef, a, b, c, seed, gen, np);
    if (rc) return rc;
    uint32_t *p = alloc_n<uint32_t>(2 * np);
    if (!p) return WTC_ERR_NOMEM;
    par_chunks(np, 1 << 18, [&](uint64_t b0, uint64_t e0, int) {
        for (uint64_t i = b0; i < e0; i++) gen(i, p[2 * i], p[2 * i + 1]);
    });
    *pairs = p;
    *npairs = np;
    return WTC_OK;
}

int wtc_gen_chung_lu(uint32_t n, double gamma, double wmin, double wmax, uint64_t target, uint64_t seed,
                     wtc_big_graph **out) {
    if (!out || n < 2 || wmin <= 0 || wmax < wmin || gamma == 1.0) return WTC_ERR_ARG;
    if (target > (uint64_t)n * (n - 1) / 2) return WTC_ERR_ARG;
    *out = nullptr;
    std::vector<double> cum(n);
    const double e1 = 1.0 - gamma, lo = std::pow(wmin, e1), hi = std::pow(wmax, e1);
    double acc = 0;
    for (uint32_t i = 0; i < n; i++) {
        double u = rng_unit(rng(seed, 1, i));
        acc += std::pow(lo + u * (hi - lo), 1.0 / e1);
        cum[i] = acc;
    }
    auto pick = [&](uint64_t x) -> uint32_t {
        double t = rng_unit(x) * acc;
        uint32_t k = (uint32_t)(std::upper_bound(cum.begin(), cum.end(), t) - cum.begin());
        return k < n ? k : n - 1;
    };
    std::unordered_set<uint64_t> seen;
    seen.reserve(target * 2);
    std::vector<uint32_t> pairs;
    const uint64_t max_draws = target * 64 + 1024;
    for (uint64_t j = 0; seen.size() < target; j++) {
        if (j >= max_draws) return WTC_ERR_ARG;
        uint32_t s = pick(rng(seed, 2, 2 * j)), d = pick(rng(seed, 2, 2 * j + 1));
        if (s == d) continue;
        seen.insert(((uint64_t)std::min(s, d) << 32) | std::max(s, d));
        pairs.push_back(s);
        pairs.push_back(d);
    }
    return wtc_graph_from_pairs(pairs.data(), pairs.size() / 2, out);
}

int wtc_gen_erdos_renyi(uint32_t n, double p, uint64_t seed, wtc_big_graph **out) {
    if (!out || n == 0 || !(p >= 0.0) || p > 1.0) return WTC_ERR_ARG;
    *out = nullptr;
    std::vector<uint32_t> pairs;
    if (p > 0.0) {
        // Batagelj & Brandes (2005), geometric skipping over the lower triangle.
        const double lp = std::log1p(-p);
        int64_t v = 1, w = -1;
        uint64_t j = 0;
        while (v < (int64_t)n) {
            double r = rng_unit(rng(seed, 3, j++));
            int64_t skip = (p >= 1.0) ? 0 : (int64_t)std::floor(std::log1p(-r) / lp);
            w += 1 + skip;
            while (w >= v && v < (int64_t)n) {
                w -= v;
                v++;
            }
            if (v < (int64_t)n) {
                pairs.push_back((uint32_t)v);
                pairs.push_back((uint32_t)w);
            }
        }
    }
    return wtc_graph_from_pairs(pairs.data(), pairs.size() / 2, out);
}

int wtc_gen_barabasi_albert(uint32_t n, uint32_t m0, uint64_t seed, wtc_big_graph **out) {
    if (!out || m0 == 0 || m0 >= n) return WTC_ERR_ARG;
    *out = nullptr;
    std::vector<uint32_t> pairs, repeated, targets(m0);
    pairs.reserve(2ull * (n - m0) * m0);
    repeated.reserve(2ull * (n - m0) * m0);
    for (uint32_t i = 0; i < m0; i++) targets[i] = i;
    uint64_t draw = 0;
    for (uint32_t v = m0; v < n; v++) {
        for (uint32_t t : targets) {
            pairs.push_back(v);
            pairs.push_back(t);
            repeated.push_back(t);
            repeated.push_back(v);
        }
        // m0 distinct targets drawn proportional to degree
        targets.clear();
        while (targets.size() < m0) {
            uint32_t c = repeated[rng(seed, 4, draw++) % repeated.size()];
            if (std::find(targets.begin(), targets.end(), c) == targets.end()) targets.push_back(c);
        }
    }
    return wtc_graph_from_pairs(pairs.data(), pairs.size() / 2, out);
}

}  // extern "C"
