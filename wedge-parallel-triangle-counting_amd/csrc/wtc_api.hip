// wtc_api.hip -- device side of the C ABI: contexts, resident graphs, the
// reference-style tc_GPU drop-ins and the batched path.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "host_util.h"
#include "wtc.h"
#include "wtc_kernels.h"
#include "wtc_launch.h"

using namespace wtc;

namespace {

thread_local std::string g_last_error;

int hip_fail(const char *what, hipError_t e) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return WTC_ERR_HIP;
}

#define HIPC(call)                                             \
    do {                                                       \
        hipError_t _e = (call);                                \
        if (_e != hipSuccess) return hip_fail(#call, _e);      \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            cap = 0;
            if (e != hipSuccess) return e;
        }
        size_t want = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

// Grow-only pinned host buffer.
struct PinBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (p) {
            (void)hipHostFree(p);
            p = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(bytes, 4096);
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

// Per-call device workspace of a graph set.
struct Workspace {
    DevBuf ws, partials, tree, bm, desc, out;
    void release() {
        ws.release();
        partials.release();
        tree.release();
        bm.release();
        desc.release();
        out.release();
    }
};

}  // namespace

struct wtc_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    Workspace work;
    PinBuf h_desc, h_out, h_stage;
    uint32_t grid_cache[2][kChunkMax + 2] = {};
};

struct wtc_dgraph {
    wtc_ctx *ctx;
    uint32_t n;
    uint64_t m;
    int ob;
    DevBuf rowPtr, colInd;
};

struct wtc_batch {
    wtc_ctx *ctx;
    uint32_t G;
    uint64_t N, M;
    int ob;
    std::vector<uint32_t> vbase, n;
    DevBuf rowPtr, colInd;
};

namespace {

int set_device(int dev) {
    int cnt = 0;
    hipError_t e = hipGetDeviceCount(&cnt);
    if (e != hipSuccess || cnt <= 0) {
        g_last_error = "no HIP device";
        return WTC_ERR_NODEV;
    }
    if (dev < 0 || dev >= cnt) return WTC_ERR_ARG;
    HIPC(hipSetDevice(dev));
    return WTC_OK;
}

uint64_t bm_words_for(uint32_t a) { return (binom2(a) + 31u) / 32u; }

int validate_oriented_serial(const wtc_big_graph *g) {
    if (!g || !g->rowPtr || (g->numEdges && !g->colInd)) return WTC_ERR_ARG;
    return validate_oriented_rows(g->numVertices, g->numEdges, g->rowPtr, g->colInd, 0, g->numVertices);
}

// Host-filled descriptor of one graph of a set.
GraphDesc make_desc(uint32_t vbase, uint32_t n, uint32_t spread, uint32_t a, uint64_t w_begin, uint64_t w_end,
                    uint64_t &bm_words, uint64_t &rows) {
    GraphDesc d{};
    d.vbase = vbase;
    d.n = n;
    d.spread = spread;
    d.a = a;
    d.bm_rows = a ? std::min(n, a) : 0u;
    d.thr = a ? vbase + (std::max(n, a) - a) : vbase + n;
    d.bm_size = binom2(a);
    d.bm_word_off = bm_words;
    d.row_base = rows;
    d.w_begin = w_begin;
    d.w_end = w_end;
    bm_words += a ? bm_words_for(a) : 0;
    rows += d.bm_rows;
    return d;
}

struct RunTimes {
    float prefix_ms = 0, bitmap_ms = 0, count_ms = 0;
};

// Runs the device pipeline over a resident graph set and returns the
// per-graph counts on the host.  `work` provides the device workspace.
int run_set(wtc_ctx *c, Workspace &work, const void *rowPtr, int ob, const uint32_t *colInd, uint64_t N,
            std::vector<GraphDesc> &hd, uint64_t bm_words, uint64_t rows, uint64_t *counts, uint64_t *wedges,
            uint32_t *grid_out, RunTimes *rt) {
    const uint32_t G = (uint32_t)hd.size();
    if (G == 0) return WTC_OK;
    if (N == 0 || N > 0xffffffffULL) return WTC_ERR_ARG;
    uint32_t max_spread = 1;
    for (const auto &d : hd) max_spread = std::max(max_spread, d.spread);
    const bool single = G == 1;
    HIPC(work.ws.ensure((N + 1) * sizeof(uint64_t)));
    HIPC(work.partials.ensure((size_t)prefix_blocks((uint32_t)N) * sizeof(uint64_t)));
    HIPC(work.tree.ensure(kTreeSize * sizeof(uint64_t)));
    HIPC(work.desc.ensure((size_t)G * sizeof(GraphDesc)));
    HIPC(work.out.ensure((size_t)(kMetaSlots + G) * sizeof(uint64_t)));
    if (bm_words) HIPC(work.bm.ensure(bm_words * sizeof(uint32_t)));
    HIPC(c->h_desc.ensure((size_t)G * sizeof(GraphDesc)));
    HIPC(c->h_out.ensure((size_t)(kMetaSlots + G) * sizeof(uint64_t)));
    std::memcpy(c->h_desc.p, hd.data(), (size_t)G * sizeof(GraphDesc));

    uint64_t *meta = work.out.as<uint64_t>();
    unsigned long long *dcounts = reinterpret_cast<unsigned long long *>(meta + kMetaSlots);
    GraphDesc *ddesc = work.desc.as<GraphDesc>();
    hipStream_t st = c->stream;

    const int oi = ob == 8 ? 1 : 0;
    const uint32_t sci = std::min<uint32_t>(max_spread, kChunkMax);
    if (!c->grid_cache[oi][sci]) HIPC(count_grid(ob, max_spread, c->num_cus, &c->grid_cache[oi][sci]));
    const uint32_t grid = c->grid_cache[oi][sci];

    HIPC(hipMemcpyAsync(ddesc, c->h_desc.p, (size_t)G * sizeof(GraphDesc), hipMemcpyHostToDevice, st));
    HIPC(hipEventRecord(c->ev[0], st));
    HIPC(launch_prefix(rowPtr, ob, (uint32_t)N, work.partials.as<uint64_t>(), work.ws.as<uint64_t>(), st));
    HIPC(launch_setup(work.ws.as<uint64_t>(), ddesc, G, single ? work.tree.as<uint64_t>() : nullptr, meta, dcounts,
                      st));
    HIPC(hipEventRecord(c->ev[1], st));
    HIPC(launch_bitmap(rowPtr, ob, colInd, ddesc, G, rows, work.bm.as<uint32_t>(), bm_words, st));
    HIPC(hipEventRecord(c->ev[2], st));
    HIPC(launch_count(rowPtr, ob, colInd, work.ws.as<uint64_t>(), single ? work.tree.as<uint64_t>() : nullptr, ddesc,
                      G, meta, work.bm.as<uint32_t>(), dcounts, max_spread, grid, st));
    HIPC(hipEventRecord(c->ev[3], st));
    HIPC(hipMemcpyAsync(c->h_out.p, meta, (size_t)(kMetaSlots + G) * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    const uint64_t *h = c->h_out.as<uint64_t>();
    for (uint32_t g = 0; g < G; g++) counts[g] = h[kMetaSlots + g];
    if (wedges) *wedges = h[kMetaWedges];
    if (grid_out) *grid_out = grid;
    if (rt) {
        HIPC(hipEventElapsedTime(&rt->prefix_ms, c->ev[0], c->ev[1]));
        HIPC(hipEventElapsedTime(&rt->bitmap_ms, c->ev[1], c->ev[2]));
        HIPC(hipEventElapsedTime(&rt->count_ms, c->ev[2], c->ev[3]));
    }
    return WTC_OK;
}

int check_args(uint32_t spread, uint32_t a) {
    if (spread == 0) return WTC_ERR_ARG;
    if (a % 32 != 0) return WTC_ERR_ARG;
    return WTC_OK;
}

// Process-wide context for the reference-style entry points (tc_GPU has no
// context argument); serialised by a mutex like the reference's globals.
std::mutex g_default_mu;
wtc_ctx *g_default = nullptr;

int default_ctx(wtc_ctx **out) {
    if (!g_default) {
        int rc = wtc_ctx_create(0, &g_default);
        if (rc) return rc;
    }
    *out = g_default;
    return WTC_OK;
}

// tc_GPU body for either offset width: per-call allocation, copy, device
// prefix/tree/bitmap, count, copy back, free (tc.cu:315-441).
template <class OffT>
int tc_gpu_impl(uint32_t n, uint64_t m, const OffT *rowPtr, const uint32_t *colInd, uint32_t spread, uint32_t a,
                wtc_gpu_time *t, uint64_t *triangles) {
    int rc = check_args(spread, a);
    if (rc) return rc;
    if (!rowPtr || (m && !colInd) || !triangles || n == 0) return WTC_ERR_ARG;
    if ((uint64_t)rowPtr[n] != m || rowPtr[0] != 0) return WTC_ERR_CSR;
    std::lock_guard<std::mutex> lk(g_default_mu);
    wtc_ctx *c;
    if ((rc = default_ctx(&c))) return rc;
    HIPC(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    Workspace work;  // freed at the end of every call, like the reference
    DevBuf drow, dcol;
    auto cleanup = [&] {
        work.release();
        drow.release();
        dcol.release();
    };
    HIPC(hipEventRecord(c->ev[4], st));
    hipError_t e;
    if ((e = drow.ensure(((size_t)n + 1) * sizeof(OffT))) != hipSuccess ||
        (e = dcol.ensure((size_t)std::max<uint64_t>(m, 1) * sizeof(uint32_t))) != hipSuccess ||
        (e = hipMemcpyAsync(drow.p, rowPtr, ((size_t)n + 1) * sizeof(OffT), hipMemcpyHostToDevice, st)) !=
            hipSuccess ||
        (m && (e = hipMemcpyAsync(dcol.p, colInd, (size_t)m * sizeof(uint32_t), hipMemcpyHostToDevice, st)) !=
                  hipSuccess)) {
        cleanup();
        return hip_fail("tc_gpu copy", e);
    }
    if ((e = hipEventRecord(c->ev[5], st)) != hipSuccess || (e = hipEventSynchronize(c->ev[5])) != hipSuccess) {
        cleanup();
        return hip_fail("tc_gpu copy sync", e);
    }
    float copy_ms = 0;
    (void)hipEventElapsedTime(&copy_ms, c->ev[4], c->ev[5]);
    uint64_t bm_words = 0, rows = 0;
    std::vector<GraphDesc> hd{make_desc(0, n, spread, a, 0, UINT64_MAX, bm_words, rows)};
    uint64_t cnt = 0;
    RunTimes rt;
    rc = run_set(c, work, drow.p, (int)sizeof(OffT), dcol.as<uint32_t>(), n, hd, bm_words, rows, &cnt, nullptr,
                 nullptr, &rt);
    cleanup();
    if (rc) return rc;
    if (t) {
        t->copy += copy_ms;
        t->exec += rt.prefix_ms + rt.bitmap_ms + rt.count_ms;
    }
    *triangles = cnt;
    return WTC_OK;
}

}  // namespace

extern "C" {

const char *wtc_last_error_detail(void) { return g_last_error.c_str(); }

int wtc_device_count(int *count) {
    if (!count) return WTC_ERR_ARG;
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *count = (e == hipSuccess) ? c : 0;
    return WTC_OK;
}

int wtc_ctx_create(int device, wtc_ctx **out) {
    if (!out) return WTC_ERR_ARG;
    *out = nullptr;
    int rc = set_device(device);
    if (rc) return rc;
    wtc_ctx *c = new (std::nothrow) wtc_ctx();
    if (!c) return WTC_ERR_NOMEM;
    c->device = device;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete c;
        return hip_fail("hipGetDeviceProperties", e);
    }
    c->num_cus = prop.multiProcessorCount;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete c;
        return hip_fail("hipStreamCreate", e);
    }
    for (auto &ev : c->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) {
            wtc_ctx_destroy(c);
            return hip_fail("hipEventCreate", e);
        }
    *out = c;
    return WTC_OK;
}

void wtc_ctx_destroy(wtc_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->work.release();
    c->h_desc.release();
    c->h_out.release();
    c->h_stage.release();
    for (auto &ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int wtc_tc_gpu(const wtc_graph *g, uint32_t spread, uint32_t a, wtc_gpu_time *t, uint64_t *triangles) {
    if (!g) return WTC_ERR_ARG;
    return tc_gpu_impl<uint32_t>(g->numVertices, g->numEdges, g->rowPtr, g->colInd, spread, a, t, triangles);
}

int wtc_tc_gpu_big(const wtc_big_graph *g, uint32_t spread, uint32_t a, wtc_gpu_time *t, uint64_t *triangles) {
    if (!g) return WTC_ERR_ARG;
    return tc_gpu_impl<uint64_t>(g->numVertices, g->numEdges, g->rowPtr, g->colInd, spread, a, t, triangles);
}

int wtc_dgraph_upload(wtc_ctx *c, const wtc_big_graph *g, uint32_t offset_bytes, wtc_dgraph **out) {
    if (!c || !g || !out || g->numVertices == 0) return WTC_ERR_ARG;
    *out = nullptr;
    int rc = wtc_validate_oriented(g);
    if (rc) return rc;
    if (offset_bytes == 0) offset_bytes = g->numEdges < 0xffffffffULL ? 4 : 8;
    if (offset_bytes != 4 && offset_bytes != 8) return WTC_ERR_ARG;
    if (offset_bytes == 4 && g->numEdges >= 0xffffffffULL) return WTC_ERR_RANGE;
    HIPC(hipSetDevice(c->device));
    wtc_dgraph *dg = new (std::nothrow) wtc_dgraph();
    if (!dg) return WTC_ERR_NOMEM;
    dg->ctx = c;
    dg->n = g->numVertices;
    dg->m = g->numEdges;
    dg->ob = (int)offset_bytes;
    const uint64_t n1 = (uint64_t)g->numVertices + 1;
    hipError_t e = dg->rowPtr.ensure(n1 * offset_bytes);
    if (e == hipSuccess) e = dg->colInd.ensure(std::max<uint64_t>(g->numEdges, 1) * sizeof(uint32_t));
    if (e == hipSuccess) {
        if (offset_bytes == 8) {
            e = hipMemcpy(dg->rowPtr.p, g->rowPtr, n1 * 8, hipMemcpyHostToDevice);
        } else {
            std::vector<uint32_t> r32(n1);
            par_chunks(n1, 1 << 20, [&](uint64_t b, uint64_t en, int) {
                for (uint64_t i = b; i < en; i++) r32[i] = (uint32_t)g->rowPtr[i];
            });
            e = hipMemcpy(dg->rowPtr.p, r32.data(), n1 * 4, hipMemcpyHostToDevice);
        }
    }
    if (e == hipSuccess && g->numEdges)
        e = hipMemcpy(dg->colInd.p, g->colInd, g->numEdges * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        wtc_dgraph_free(dg);
        return hip_fail("dgraph upload", e);
    }
    *out = dg;
    return WTC_OK;
}

void wtc_dgraph_free(wtc_dgraph *dg) {
    if (!dg) return;
    (void)hipSetDevice(dg->ctx->device);
    dg->rowPtr.release();
    dg->colInd.release();
    delete dg;
}

uint64_t wtc_dgraph_bytes(const wtc_dgraph *dg) {
    return dg ? ((uint64_t)dg->n + 1) * dg->ob + dg->m * 4u : 0;
}

int wtc_dgraph_count(wtc_ctx *c, wtc_dgraph *dg, uint32_t spread, uint32_t a, uint64_t w_begin, uint64_t w_end,
                     uint64_t *triangles, wtc_count_stats *st) {
    if (!c || !dg || !triangles) return WTC_ERR_ARG;
    int rc = check_args(spread, a);
    if (rc) return rc;
    HIPC(hipSetDevice(c->device));
    uint64_t bm_words = 0, rows = 0;
    std::vector<GraphDesc> hd{make_desc(0, dg->n, spread, a, w_begin, w_end, bm_words, rows)};
    RunTimes rt;
    uint64_t W = 0;
    uint32_t grid = 0;
    auto t0 = std::chrono::steady_clock::now();
    rc = run_set(c, c->work, dg->rowPtr.p, dg->ob, dg->colInd.as<uint32_t>(), dg->n, hd, bm_words, rows, triangles,
                 &W, &grid, &rt);
    auto t1 = std::chrono::steady_clock::now();
    if (rc) return rc;
    if (st) {
        st->prefix_ms = rt.prefix_ms;
        st->bitmap_ms = rt.bitmap_ms;
        st->count_ms = rt.count_ms;
        st->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->wedges = W;
        st->blocks = grid;
        st->launches = 1;
    }
    return WTC_OK;
}

// Copies the wedge prefix and search tree left by the last count on this
// context (diagnostics for the parity tests).
int wtc_ctx_debug_prefix(wtc_ctx *c, uint64_t n, uint64_t *wedge_sum, uint64_t *tree) {
    if (!c) return WTC_ERR_ARG;
    HIPC(hipSetDevice(c->device));
    if (wedge_sum) {
        if (c->work.ws.cap < (n + 1) * sizeof(uint64_t)) return WTC_ERR_ARG;
        HIPC(hipMemcpy(wedge_sum, c->work.ws.p, (n + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
    }
    if (tree) HIPC(hipMemcpy(tree, c->work.tree.p, kTreeSize * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return WTC_OK;
}

int wtc_debug_decode(wtc_ctx *c, const uint32_t *d, const uint64_t *iv, uint64_t k, uint32_t *w_out,
                     uint32_t *u_out) {
    if (!c || (k && (!d || !iv || !w_out || !u_out))) return WTC_ERR_ARG;
    for (uint64_t i = 0; i < k; i++)
        if (d[i] < 2 || iv[i] >= binom2(d[i])) return WTC_ERR_ARG;
    if (k == 0) return WTC_OK;
    HIPC(hipSetDevice(c->device));
    DevBuf b;
    hipError_t e = b.ensure(k * (4 + 8 + 4 + 4));
    if (e != hipSuccess) return hip_fail("debug_decode alloc", e);
    char *p = b.as<char>();
    uint64_t *di = reinterpret_cast<uint64_t *>(p);  // 8-byte entries first, at the aligned base
    uint32_t *dd = reinterpret_cast<uint32_t *>(p + 8 * k), *dw = dd + k, *du = dw + k;
    if ((e = hipMemcpyAsync(dd, d, k * 4, hipMemcpyHostToDevice, c->stream)) == hipSuccess &&
        (e = hipMemcpyAsync(di, iv, k * 8, hipMemcpyHostToDevice, c->stream)) == hipSuccess &&
        (e = launch_debug_decode(dd, di, k, dw, du, c->stream)) == hipSuccess &&
        (e = hipMemcpyAsync(w_out, dw, k * 4, hipMemcpyDeviceToHost, c->stream)) == hipSuccess &&
        (e = hipMemcpyAsync(u_out, du, k * 4, hipMemcpyDeviceToHost, c->stream)) == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    b.release();
    return e == hipSuccess ? WTC_OK : hip_fail("debug_decode", e);
}

int wtc_ctx_debug_bitmap(wtc_ctx *c, uint64_t words, uint32_t *out) {
    if (!c || !out) return WTC_ERR_ARG;
    if (c->work.bm.cap < words * sizeof(uint32_t)) return WTC_ERR_ARG;
    HIPC(hipSetDevice(c->device));
    HIPC(hipMemcpy(out, c->work.bm.p, words * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return WTC_OK;
}

// ------------------------------------------------------------------ batches

void wtc_autotune(uint32_t n, uint64_t m, uint64_t wedges, uint32_t *spread, uint32_t *a) {
    (void)m;
    uint32_t s, aa;
    if (wedges < (1u << 16)) s = 2;
    else if (wedges < (1u << 22)) s = 4;
    else s = 8;
    if (n <= 1024) aa = (n + 31) & ~31u;
    else if (n <= 16384) aa = 1024;
    else aa = 2048;
    if (spread) *spread = s;
    if (a) *a = aa;
}

int wtc_batch_upload(wtc_ctx *c, const wtc_big_graph *const *graphs, size_t k, wtc_batch **out) {
    if (!c || !out || (k && !graphs) || k > 0xffffffffULL) return WTC_ERR_ARG;
    *out = nullptr;
    uint64_t N = 0, M = 0;
    std::vector<uint64_t> ebase(k);
    wtc_batch *b = new (std::nothrow) wtc_batch();
    if (!b) return WTC_ERR_NOMEM;
    b->ctx = c;
    b->G = (uint32_t)k;
    b->vbase.resize(k);
    b->n.resize(k);
    for (size_t g = 0; g < k; g++) {
        if (!graphs[g] || graphs[g]->numVertices == 0) {
            delete b;
            return WTC_ERR_ARG;
        }
        b->vbase[g] = (uint32_t)N;
        b->n[g] = graphs[g]->numVertices;
        ebase[g] = M;
        N += graphs[g]->numVertices;
        M += graphs[g]->numEdges;
        if (N > 0xfffffffeULL) {
            delete b;
            return WTC_ERR_RANGE;
        }
    }
    std::atomic<int> bad{0};
    par_chunks(k, 1, [&](uint64_t g0, uint64_t g1, int) {
        for (uint64_t g = g0; g < g1; g++)
            if (validate_oriented_serial(graphs[g]) != WTC_OK) bad = 1;
    });
    if (bad) {
        delete b;
        return WTC_ERR_CSR;
    }
    b->N = N;
    b->M = M;
    b->ob = M < 0xffffffffULL ? 4 : 8;
    hipError_t e = hipSetDevice(c->device);
    const size_t rbytes = (N + 1) * (size_t)b->ob, cbytes = std::max<uint64_t>(M, 1) * 4;
    if (e == hipSuccess) e = b->rowPtr.ensure(rbytes);
    if (e == hipSuccess) e = b->colInd.ensure(cbytes);
    if (e == hipSuccess) e = c->h_stage.ensure(rbytes + cbytes);
    if (e != hipSuccess) {
        wtc_batch_free(b);
        return hip_fail("batch alloc", e);
    }
    // pack the disjoint union into pinned staging: ids and offsets shifted
    char *stage = c->h_stage.as<char>();
    uint32_t *sc = reinterpret_cast<uint32_t *>(stage + rbytes);
    const int ob = b->ob;
    par_chunks(k, 1, [&](uint64_t g0, uint64_t g1, int) {
        for (uint64_t g = g0; g < g1; g++) {
            const wtc_big_graph *gr = graphs[g];
            const uint32_t vb = b->vbase[g];
            const uint64_t eb = ebase[g];
            for (uint64_t v = 0; v < gr->numVertices; v++) {
                const uint64_t off = eb + gr->rowPtr[v];
                if (ob == 8) reinterpret_cast<uint64_t *>(stage)[vb + v] = off;
                else reinterpret_cast<uint32_t *>(stage)[vb + v] = (uint32_t)off;
            }
            for (uint64_t j = 0; j < gr->numEdges; j++) sc[eb + j] = gr->colInd[j] + vb;
        }
    });
    if (ob == 8) reinterpret_cast<uint64_t *>(stage)[N] = M;
    else reinterpret_cast<uint32_t *>(stage)[N] = (uint32_t)M;
    e = hipMemcpyAsync(b->rowPtr.p, stage, rbytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && M) e = hipMemcpyAsync(b->colInd.p, sc, M * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        wtc_batch_free(b);
        return hip_fail("batch upload", e);
    }
    *out = b;
    return WTC_OK;
}

void wtc_batch_free(wtc_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    b->rowPtr.release();
    b->colInd.release();
    delete b;
}

int wtc_batch_count(wtc_ctx *c, wtc_batch *b, const uint32_t *spread, const uint32_t *a, uint64_t *triangles,
                    wtc_count_stats *st) {
    if (!c || !b || !triangles) return WTC_ERR_ARG;
    if (b->G == 0) return WTC_OK;
    HIPC(hipSetDevice(c->device));
    std::vector<GraphDesc> hd(b->G);
    uint64_t bm_words = 0, rows = 0;
    for (uint32_t g = 0; g < b->G; g++) {
        uint32_t s = spread ? spread[g] : 0, aa = a ? a[g] : 0;
        if (s == 0) wtc_autotune(b->n[g], 0, 0, &s, a ? nullptr : &aa);
        if (s == 0 || aa % 32 != 0) return WTC_ERR_ARG;
        hd[g] = make_desc(b->vbase[g], b->n[g], s, aa, 0, UINT64_MAX, bm_words, rows);
    }
    RunTimes rt;
    uint64_t W = 0;
    uint32_t grid = 0;
    auto t0 = std::chrono::steady_clock::now();
    int rc = run_set(c, c->work, b->rowPtr.p, b->ob, b->colInd.as<uint32_t>(), b->N, hd, bm_words, rows, triangles,
                     &W, &grid, &rt);
    auto t1 = std::chrono::steady_clock::now();
    if (rc) return rc;
    if (st) {
        st->prefix_ms = rt.prefix_ms;
        st->bitmap_ms = rt.bitmap_ms;
        st->count_ms = rt.count_ms;
        st->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->wedges = W;
        st->blocks = grid;
        st->launches = 1;
    }
    return WTC_OK;
}

int wtc_batch_count_host(wtc_ctx *c, const wtc_big_graph *const *graphs, size_t k, const uint32_t *spread,
                         const uint32_t *a, uint64_t *triangles, wtc_count_stats *st) {
    auto t0 = std::chrono::steady_clock::now();
    wtc_batch *b = nullptr;
    int rc = wtc_batch_upload(c, graphs, k, &b);
    if (rc) return rc;
    rc = wtc_batch_count(c, b, spread, a, triangles, st);
    wtc_batch_free(b);
    auto t1 = std::chrono::steady_clock::now();
    if (st && rc == WTC_OK) st->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return rc;
}

}  // extern "C"
