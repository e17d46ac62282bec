// wtc_alt.hip -- the reference's alternative counts from experiments/, on the
// same resident CSR and top-a bitmap as the wedge-parallel count.
//
// 1. The arrow-wedge count of experiments/tc_arrow.cu (and its 64-bit twin
// experiments/tc_big_graphs_arrow.cu): every directed path
// v -> w -> u of the oriented CSR is a wedge, closed when u is in adj+(v).
// Each triangle v < w < u (in rank order) is one closed arrow wedge, so the
// count equals the wedge-parallel count; the wedge set (and so the work) is
// different: sum over edges (v,w) of d+(w) instead of sum over v of C(d+(v),2).
//
// Device pipeline on the context's stream:
//   k_arrow_degrees / k_arrow_single_rows
//       x[e] = d+(colInd[e]), then 0 for the edge of every row with d+ = 1
//       (the host prefix of tc_arrow.cu:345-368 gives such rows no wedges)
//   k_arrow_partials / k_scan_partials / k_arrow_scan
//       vi = exclusive prefix of x over the m edges, in place (vi[m] = total)
//   k_bitmap_clear / k_bitmap_build
//       the top-a bitmap, shared with the wedge-parallel count
//   k_arrow_vertex_prefix
//       vp[v] = vi[rowPtr[v]], the first arrow wedge of every source
//   k_count_arrow
//       persistent grid; lane-contiguous chunks of `spread` wedges, located
//       by two searches (source over vp, edge over vi within the source's
//       row) and then stepped;
//       u probed in adj+(v) by the bitmap for v in the top-a ranks, else by a
//       binary search of the row.
//
// 2. The vertex- and edge-parallel baselines of experiments/tc_vertex.cu and
// experiments/tc_edge.cu (k_count_vertex / k_count_edge): one lane per vertex,
// or per oriented edge (its source found by a search of rowPtr), with the
// reference's three wedge styles per edge (v,w):
//   outgoing  u over adj+(v) after w, probed in adj+(w)
//   arrow     u over adj+(w), probed in adj+(v)
//   mixed     whichever of the two lists is shorter is walked, the other probed
// Their work per lane is a whole vertex or edge, so they are load-imbalanced by
// design; they are kept as the baselines the wedge-parallel count replaces.
#include <chrono>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "wtc_dev.h"
#include "wtc_internal.h"
#include "wtc_launch.h"

namespace wtc {
namespace {

template <class OffT>
__global__ __launch_bounds__(kBlock) void k_arrow_degrees(const void *rp, const uint32_t *__restrict__ colInd,
                                                          uint64_t m, uint64_t *__restrict__ x) {
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(rp);
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e < m; e += stride) {
        const uint32_t w = colInd[e];
        x[e] = (uint64_t)(rowPtr[w + 1] - rowPtr[w]);
    }
}

template <class OffT>
__global__ __launch_bounds__(kBlock) void k_arrow_single_rows(const void *rp, uint32_t n, uint64_t *__restrict__ x) {
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(rp);
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
        const uint64_t vb = rowPtr[v];
        if ((uint64_t)rowPtr[v + 1] - vb == 1) x[vb] = 0;
    }
}

// Block b sums x over [b*kScanChunk, (b+1)*kScanChunk) of [0, m).
__global__ __launch_bounds__(kBlock) void k_arrow_partials(const uint64_t *__restrict__ x, uint64_t m,
                                                           uint64_t *__restrict__ partials) {
    __shared__ uint64_t s_tot[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    uint64_t acc = 0;
#pragma unroll 4
    for (int k = 0; k < kScanItems; k++) {
        const uint64_t e = base + (uint64_t)k * kBlock + threadIdx.x;
        if (e < m) acc += x[e];
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) s_tot[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = s_tot[0] + s_tot[1] + s_tot[2] + s_tot[3];
}

// In place: each thread reads its x[e] before any thread of the block writes
// (block_excl_scan synchronises), and blocks own disjoint ranges, so
// vi[e] = partials[b] + (exclusive scan inside the block); the thread of the
// last edge also writes vi[m].
__global__ __launch_bounds__(kBlock) void k_arrow_scan(uint64_t *__restrict__ xv, uint64_t m,
                                                       const uint64_t *__restrict__ partials) {
    __shared__ uint64_t s_tot[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    uint64_t carry = partials[blockIdx.x];
    for (int k = 0; k < kScanItems; k++) {
        const uint64_t e = base + (uint64_t)k * kBlock + threadIdx.x;
        const uint64_t x = e < m ? xv[e] : 0;
        uint64_t tot;
        const uint64_t ex = block_excl_scan(x, s_tot, tot);
        if (e < m) xv[e] = carry + ex;
        if (e + 1 == m) xv[m] = carry + ex + x;
        carry += tot;
    }
}

// vp[v] = vi[rowPtr[v]] for v in [0, n]: the first arrow wedge of each
// source vertex (the reference's wedge_vp, tc_arrow.cu:362-366).
template <class OffT>
__global__ __launch_bounds__(kBlock) void k_arrow_vertex_prefix(const void *rp, const uint64_t *__restrict__ vi,
                                                                uint32_t n, uint64_t *__restrict__ vp) {
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(rp);
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v <= n; v += stride) vp[v] = vi[rowPtr[v]];
}

// Rightmost index i in [0, hi] with a[i] <= t, given a[0] <= t.
template <class T>
__device__ __forceinline__ uint64_t rightmost_le64(const T *__restrict__ a, uint64_t hi, uint64_t t) {
    uint64_t s = 1, e = hi + 1;
    while (s < e) {
        const uint64_t mid = s + ((e - s) >> 1);
        if ((uint64_t)a[mid] <= t) s = mid + 1;
        else e = mid;
    }
    return s - 1;
}

// Lower bound of u in the ascending list[lo, end): first position whose id
// is >= u, or end.
__device__ __forceinline__ uint64_t lower_bound(const uint32_t *__restrict__ list, uint64_t lo, uint64_t end,
                                                uint32_t u) {
    uint64_t hi = end;
    while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        if (list[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Chunk c covers arrow wedges [c*spread, min((c+1)*spread, A)); thread t of
// the grid takes chunks t, t + T, ...  Adjacent lanes take adjacent chunks,
// so a wave mostly works inside one row.
template <class OffT>
__global__ __launch_bounds__(kBlock) void k_count_arrow(const void *rp, const uint32_t *__restrict__ colInd,
                                                        uint32_t n, uint64_t m,
                                                        const uint64_t *__restrict__ vi,
                                                        const uint64_t *__restrict__ vp,
                                                        const uint32_t *__restrict__ bitmap, uint32_t a32,
                                                        uint32_t thr, uint32_t spread,
                                                        unsigned long long *__restrict__ count) {
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(rp);
    const uint64_t A = vi[m];
    const uint64_t nchunks = (A + spread - 1) / spread;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    uint64_t total = 0;
    for (uint64_t c = (uint64_t)blockIdx.x * kBlock + threadIdx.x; c < nchunks; c += stride) {
        uint64_t i = c * spread;
        const uint64_t iend = i + spread < A ? i + spread : A;
        // the source first (vp[0] = 0 <= i < vp[n] = A, so vp[v] <= i < vp[v+1]
        // and row v holds at least two edges), then its edge e in [vb, ve)
        // with vi[e] <= i < vi[e+1] (vi[vb] = vp[v] <= i < vi[ve] = vp[v+1])
        uint32_t v = (uint32_t)rightmost_le64(vp, n, i);
        uint64_t vb = rowPtr[v], ve = rowPtr[v + 1];
        uint64_t e = vb + rightmost_le64(vi + vb, ve - 1 - vb, i);
        uint64_t eb = vi[e], ee = vi[e + 1];
        uint64_t wb = rowPtr[colInd[e]];
        // the u of one edge ascend, so each search starts at the previous
        // lower bound in adj+(v) (reset when the edge changes)
        uint64_t lo = vb;
        for (; i < iend; i++) {
            if (i >= ee) {  // next edge with wedges (i < A keeps e + 1 <= m)
                do {
                    e++;
                    eb = ee;
                    ee = vi[e + 1];
                } while (i >= ee);
                while (e >= ve) {  // next source row holding edge e
                    v++;
                    vb = ve;
                    ve = rowPtr[v + 1];
                }
                wb = rowPtr[colInd[e]];
                lo = vb;
            }
            const uint32_t u = colInd[wb + (i - eb)];
            if (v >= thr) {
                total += bm_test(bitmap, v, u, thr, a32);
            } else {
                lo = lower_bound(colInd, lo, ve, u);
                total += (lo < ve && colInd[lo] == u) ? 1u : 0u;
            }
        }
    }
    total = wave_sum(total);
    if ((threadIdx.x & (kWave - 1)) == 0 && total) atomicAdd(count, (unsigned long long)total);
}

// Triangle-closing probe of u: the bitmap row of x when x is in the top-a
// ranks, else a binary search of list[lo, hi).
struct Probe {
    const uint32_t *bitmap;
    uint32_t a32;
    uint32_t thr, n;
    __device__ __forceinline__ uint32_t has(const uint32_t *__restrict__ colInd, uint32_t x, uint64_t lo,
                                            uint64_t hi, uint32_t u) const {
        if (x >= thr) return bm_test(bitmap, x, u, thr, a32);
        const uint64_t p = lower_bound(colInd, lo, hi, u);
        return (p < hi && colInd[p] == u) ? 1u : 0u;
    }
};

// Closed wedges of edge i = (v, w) of row [vb, ve) in wedge style `Style`
// (tc_vertex.cu:113-274 / tc_edge.cu:179-352, per edge).
template <class OffT, int Style>
__device__ __forceinline__ uint64_t edge_closed(const OffT *__restrict__ rowPtr, const uint32_t *__restrict__ colInd,
                                                const Probe &pr, uint32_t v, uint64_t vb, uint64_t ve, uint64_t i) {
    const uint32_t w = colInd[i];
    const uint64_t wb = rowPtr[w], we = rowPtr[w + 1];
    bool walk_v = Style == 0;
    if (Style == 2) walk_v = ve - (i + 1) <= we - wb;
    uint64_t c = 0;
    if (walk_v) {  // u after w in adj+(v), probed in adj+(w)
        for (uint64_t j = i + 1; j < ve; j++) c += pr.has(colInd, w, wb, we, colInd[j]);
    } else {  // u in adj+(w), probed in adj+(v) (after w for the mixed style)
        const uint64_t lo = Style == 2 ? i + 1 : vb;
        for (uint64_t j = wb; j < we; j++) c += pr.has(colInd, v, lo, ve, colInd[j]);
    }
    return c;
}

template <class OffT, int Style>
__global__ __launch_bounds__(kBlock) void k_count_vertex(const void *rp, const uint32_t *__restrict__ colInd,
                                                         uint32_t n, Probe pr,
                                                         unsigned long long *__restrict__ count) {
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(rp);
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    uint64_t total = 0;
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
        const uint64_t vb = rowPtr[v], ve = rowPtr[v + 1];
        for (uint64_t i = vb; i < ve; i++) total += edge_closed<OffT, Style>(rowPtr, colInd, pr, (uint32_t)v, vb, ve, i);
    }
    total = wave_sum(total);
    if ((threadIdx.x & (kWave - 1)) == 0 && total) atomicAdd(count, (unsigned long long)total);
}

template <class OffT, int Style>
__global__ __launch_bounds__(kBlock) void k_count_edge(const void *rp, const uint32_t *__restrict__ colInd,
                                                       uint32_t n, uint64_t m, Probe pr,
                                                       unsigned long long *__restrict__ count) {
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(rp);
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    uint64_t total = 0;
    for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e < m; e += stride) {
        // rowPtr[n] = m > e: the source v < n has rowPtr[v] <= e < rowPtr[v+1]
        const uint32_t v = (uint32_t)rightmost_le64(rowPtr, n, e);
        total += edge_closed<OffT, Style>(rowPtr, colInd, pr, v, rowPtr[v], rowPtr[v + 1], e);
    }
    total = wave_sum(total);
    if ((threadIdx.x & (kWave - 1)) == 0 && total) atomicAdd(count, (unsigned long long)total);
}

template <int Style>
hipError_t launch_baseline(int ob, int edge_parallel, dim3 grid, hipStream_t st, const void *rp,
                           const uint32_t *colInd, uint32_t n, uint64_t m, const Probe &pr,
                           unsigned long long *count) {
    if (edge_parallel) {
        if (ob == 8) k_count_edge<uint64_t, Style><<<grid, dim3(kBlock), 0, st>>>(rp, colInd, n, m, pr, count);
        else k_count_edge<uint32_t, Style><<<grid, dim3(kBlock), 0, st>>>(rp, colInd, n, m, pr, count);
    } else {
        if (ob == 8) k_count_vertex<uint64_t, Style><<<grid, dim3(kBlock), 0, st>>>(rp, colInd, n, pr, count);
        else k_count_vertex<uint32_t, Style><<<grid, dim3(kBlock), 0, st>>>(rp, colInd, n, pr, count);
    }
    return hipGetLastError();
}

#define WTC_ARROW_LAUNCH(ob, KERNEL, GRID, STREAM, ...)                        \
    [&]() -> hipError_t {                                                      \
        if ((ob) == 8)                                                         \
            KERNEL<uint64_t><<<GRID, dim3(kBlock), 0, STREAM>>>(__VA_ARGS__);  \
        else                                                                   \
            KERNEL<uint32_t><<<GRID, dim3(kBlock), 0, STREAM>>>(__VA_ARGS__);  \
        return hipGetLastError();                                              \
    }()

// Descriptor of a resident graph with its top-a bitmap in the slot's
// workspace; the descriptor is staged in the slot's pinned buffer for
// enqueue_bitmap.
int prepare_bitmap(Slot &s, const wtc_dgraph *dg, uint32_t a, GraphDesc &d, uint64_t &bm_words) {
    d = make_desc(dg->rowPtr.p, dg->colInd.as<uint32_t>(), dg->n, 1, a, 0, UINT64_MAX);
    bm_words = a ? bm_words_for(a) : 0;
    if (!bm_words) return WTC_OK;
    WTC_HIPC(s.bm.ensure(bm_words * sizeof(uint32_t)));
    d.bitmap = s.bm.as<uint32_t>();
    WTC_HIPC(s.desc.ensure(sizeof(GraphDesc)));
    WTC_HIPC(s.h_desc.ensure(sizeof(GraphDesc)));
    std::memcpy(s.h_desc.p, &d, sizeof d);
    return WTC_OK;
}

// Clears and builds the bitmap prepared by prepare_bitmap on the slot's stream.
int enqueue_bitmap(Slot &s, const GraphDesc &d, int ob, uint64_t bm_words) {
    if (!bm_words) return WTC_OK;
    WTC_HIPC(hipMemcpyAsync(s.desc.p, s.h_desc.p, sizeof(GraphDesc), hipMemcpyHostToDevice, s.stream));
    WTC_HIPC(launch_bitmap(s.desc.as<GraphDesc>(), 1, ob, d.bm_rows, bm_words, s.stream));
    return WTC_OK;
}

uint32_t stride_grid(uint64_t items, uint32_t cap) {
    const uint64_t b = (items + kBlock - 1) / kBlock;
    return (uint32_t)(b < 1 ? 1 : b < cap ? b : cap);
}

}  // namespace
}  // namespace wtc

using namespace wtc;

extern "C" {

int wtc_dgraph_count_arrow(wtc_ctx *c, wtc_dgraph *dg, uint32_t spread, uint32_t a, uint64_t *triangles,
                           uint64_t *arrow_wedges, wtc_count_stats *st) {
    if (!c || !dg || !triangles) return WTC_ERR_ARG;
    int rc = check_args(spread, a);
    if (rc) return rc;
    if (dg->m > 0xffffffffULL * kScanChunk) return WTC_ERR_RANGE;
    WTC_HIPC(hipSetDevice(c->device));
    auto t0 = std::chrono::steady_clock::now();
    Slot &s = c->slot;
    hipStream_t stream = s.stream;
    const int ob = dg->ob;
    const uint32_t n = dg->n;
    const uint64_t m = dg->m;
    const uint32_t nb = (uint32_t)((m + kScanChunk - 1) / kScanChunk);
    WTC_HIPC(s.ws.ensure((m + 1 + (uint64_t)n + 1) * sizeof(uint64_t)));  // vi, then vp
    WTC_HIPC(s.partials.ensure(std::max<uint32_t>(nb, 1) * sizeof(uint64_t)));
    WTC_HIPC(s.out.ensure((size_t)kMetaSlots * sizeof(uint64_t)));
    WTC_HIPC(s.h_out.ensure((size_t)kMetaSlots * sizeof(uint64_t)));
    GraphDesc d;
    uint64_t bm_words = 0;
    if ((rc = prepare_bitmap(s, dg, a, d, bm_words))) return rc;
    uint64_t *vi = s.ws.as<uint64_t>();
    uint64_t *vp = vi + m + 1;
    uint64_t *out = s.out.as<uint64_t>();
    unsigned long long *dcount = reinterpret_cast<unsigned long long *>(out);
    const uint32_t cap = (uint32_t)std::max(c->tune.num_cus, 1) * 16u;

    WTC_HIPC(hipEventRecord(s.ev[0], stream));
    WTC_HIPC(launch_fill(reinterpret_cast<uint32_t *>(out), 2, 0u, stream));
    if (m) {
        WTC_HIPC(WTC_ARROW_LAUNCH(ob, k_arrow_degrees, dim3(stride_grid(m, cap)), stream, dg->rowPtr.p,
                                  dg->colInd.as<uint32_t>(), m, vi));
        WTC_HIPC(WTC_ARROW_LAUNCH(ob, k_arrow_single_rows, dim3(stride_grid(n, cap)), stream, dg->rowPtr.p, n, vi));
        k_arrow_partials<<<dim3(nb), dim3(kBlock), 0, stream>>>(vi, m, s.partials.as<uint64_t>());
        WTC_HIPC(hipGetLastError());
        WTC_HIPC(launch_scan_partials(s.partials.as<uint64_t>(), nb, stream));
        k_arrow_scan<<<dim3(nb), dim3(kBlock), 0, stream>>>(vi, m, s.partials.as<uint64_t>());
        WTC_HIPC(hipGetLastError());
        WTC_HIPC(WTC_ARROW_LAUNCH(ob, k_arrow_vertex_prefix, dim3(stride_grid((uint64_t)n + 1, cap)), stream,
                                  dg->rowPtr.p, vi, n, vp));
    } else {
        WTC_HIPC(launch_fill(reinterpret_cast<uint32_t *>(vi), 2, 0u, stream));  // vi[0] = A = 0
    }
    WTC_HIPC(hipEventRecord(s.ev[1], stream));
    if ((rc = enqueue_bitmap(s, d, ob, bm_words))) return rc;
    WTC_HIPC(hipEventRecord(s.ev[2], stream));
    // chunks of `spread` wedges, one per lane of a persistent grid
    const uint32_t grid = cap * 2u;
    if (m)
        WTC_HIPC(WTC_ARROW_LAUNCH(ob, k_count_arrow, dim3(grid), stream, dg->rowPtr.p, dg->colInd.as<uint32_t>(), n,
                                  m, vi, vp, d.bitmap, d.a >> 5, d.thr, spread, dcount));
    WTC_HIPC(hipEventRecord(s.ev[3], stream));
    WTC_HIPC(hipMemcpyAsync(s.h_out.p, out, sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
    WTC_HIPC(hipMemcpyAsync(s.h_out.as<uint64_t>() + 1, vi + m, sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
    WTC_HIPC(hipStreamSynchronize(stream));
    auto t1 = std::chrono::steady_clock::now();
    const uint64_t *h = s.h_out.as<uint64_t>();
    *triangles = h[0];
    if (arrow_wedges) *arrow_wedges = h[1];
    if (st) {
        float p = 0, b = 0, k = 0;
        WTC_HIPC(hipEventElapsedTime(&p, s.ev[0], s.ev[1]));
        WTC_HIPC(hipEventElapsedTime(&b, s.ev[1], s.ev[2]));
        WTC_HIPC(hipEventElapsedTime(&k, s.ev[2], s.ev[3]));
        *st = wtc_count_stats{};
        st->prefix_ms = p;
        st->bitmap_ms = b;
        st->count_ms = k;
        st->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->wedges = h[1];
        st->blocks = m ? grid : 0;
        st->launches = m ? 1 : 0;
    }
    return WTC_OK;
}

int wtc_dgraph_count_baseline(wtc_ctx *c, wtc_dgraph *dg, int parallelism, int style, uint32_t a,
                              uint64_t *triangles, wtc_count_stats *st) {
    if (!c || !dg || !triangles) return WTC_ERR_ARG;
    if (parallelism != WTC_PARALLEL_VERTEX && parallelism != WTC_PARALLEL_EDGE) return WTC_ERR_ARG;
    if (style < WTC_WEDGE_OUTGOING || style > WTC_WEDGE_MIXED) return WTC_ERR_ARG;
    int rc = check_args(1, a);
    if (rc) return rc;
    WTC_HIPC(hipSetDevice(c->device));
    auto t0 = std::chrono::steady_clock::now();
    Slot &s = c->slot;
    WTC_HIPC(s.out.ensure((size_t)kMetaSlots * sizeof(uint64_t)));
    WTC_HIPC(s.h_out.ensure((size_t)kMetaSlots * sizeof(uint64_t)));
    GraphDesc d;
    uint64_t bm_words = 0;
    if ((rc = prepare_bitmap(s, dg, a, d, bm_words))) return rc;
    unsigned long long *dcount = reinterpret_cast<unsigned long long *>(s.out.p);
    const uint64_t items = parallelism == WTC_PARALLEL_EDGE ? dg->m : dg->n;
    const dim3 grid(stride_grid(items, (uint32_t)std::max(c->tune.num_cus, 1) * 32u));
    const Probe pr{d.bitmap, d.a >> 5, d.thr, dg->n};
    WTC_HIPC(hipEventRecord(s.ev[0], s.stream));
    WTC_HIPC(launch_fill(reinterpret_cast<uint32_t *>(s.out.p), 2, 0u, s.stream));
    WTC_HIPC(hipEventRecord(s.ev[1], s.stream));
    if ((rc = enqueue_bitmap(s, d, dg->ob, bm_words))) return rc;
    WTC_HIPC(hipEventRecord(s.ev[2], s.stream));
    if (dg->m) {
        const int ep = parallelism == WTC_PARALLEL_EDGE;
        const void *rp = dg->rowPtr.p;
        const uint32_t *ci = dg->colInd.as<uint32_t>();
        if (style == WTC_WEDGE_OUTGOING)
            WTC_HIPC(launch_baseline<0>(dg->ob, ep, grid, s.stream, rp, ci, dg->n, dg->m, pr, dcount));
        else if (style == WTC_WEDGE_ARROW)
            WTC_HIPC(launch_baseline<1>(dg->ob, ep, grid, s.stream, rp, ci, dg->n, dg->m, pr, dcount));
        else
            WTC_HIPC(launch_baseline<2>(dg->ob, ep, grid, s.stream, rp, ci, dg->n, dg->m, pr, dcount));
    }
    WTC_HIPC(hipEventRecord(s.ev[3], s.stream));
    WTC_HIPC(hipMemcpyAsync(s.h_out.p, s.out.p, sizeof(uint64_t), hipMemcpyDeviceToHost, s.stream));
    WTC_HIPC(hipStreamSynchronize(s.stream));
    auto t1 = std::chrono::steady_clock::now();
    *triangles = s.h_out.as<uint64_t>()[0];
    if (st) {
        float b = 0, k = 0;
        WTC_HIPC(hipEventElapsedTime(&b, s.ev[1], s.ev[2]));
        WTC_HIPC(hipEventElapsedTime(&k, s.ev[2], s.ev[3]));
        *st = wtc_count_stats{};
        st->bitmap_ms = b;
        st->count_ms = k;
        st->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->blocks = dg->m ? grid.x : 0;
        st->launches = dg->m ? 1 : 0;
    }
    return WTC_OK;
}

int wtc_tc_arrow_gpu_big(const wtc_big_graph *g, uint32_t spread, uint32_t a, wtc_gpu_time *t,
                         uint64_t *triangles) {
    if (!g || !triangles) return WTC_ERR_ARG;
    int rc = check_args(spread, a);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(default_ctx_mutex());
    wtc_ctx *c;
    if ((rc = default_context(&c))) return rc;
    auto t0 = std::chrono::steady_clock::now();
    wtc_dgraph *dg = nullptr;
    if ((rc = wtc_dgraph_upload(c, g, 0, &dg))) return rc;
    auto t1 = std::chrono::steady_clock::now();
    wtc_count_stats st{};
    rc = wtc_dgraph_count_arrow(c, dg, spread, a, triangles, nullptr, &st);
    wtc_dgraph_free(dg);
    c->slot.release_device();  // per-call buffers, like the reference
    if (rc) return rc;
    if (t) {
        t->copy += std::chrono::duration<double, std::milli>(t1 - t0).count();
        t->exec += st.prefix_ms + st.bitmap_ms + st.count_ms;
    }
    return WTC_OK;
}


// Per-call form shared by the tc_vertex_GPU / tc_edge_GPU drop-ins: upload,
// count, free, with the copy and device times accumulated like GPU_time.
static int baseline_call(const wtc_big_graph *g, int parallelism, uint32_t a, int style, wtc_gpu_time *t,
                         uint64_t *triangles) {
    if (!g || !triangles) return WTC_ERR_ARG;
    std::lock_guard<std::mutex> lk(default_ctx_mutex());
    wtc_ctx *c;
    int rc = default_context(&c);
    if (rc) return rc;
    auto t0 = std::chrono::steady_clock::now();
    wtc_dgraph *dg = nullptr;
    if ((rc = wtc_dgraph_upload(c, g, 0, &dg))) return rc;
    auto t1 = std::chrono::steady_clock::now();
    wtc_count_stats st{};
    rc = wtc_dgraph_count_baseline(c, dg, parallelism, style, a, triangles, &st);
    wtc_dgraph_free(dg);
    c->slot.release_device();
    if (rc) return rc;
    if (t) {
        t->copy += std::chrono::duration<double, std::milli>(t1 - t0).count();
        t->exec += st.bitmap_ms + st.count_ms;
    }
    return WTC_OK;
}

// 32-bit offsets widened on the host (the upload narrows them again).
static int widen(const wtc_graph *g, std::vector<uint64_t> &rp, wtc_big_graph &b) {
    if (!g || !g->rowPtr || g->numVertices == 0) return WTC_ERR_ARG;
    try {
        rp.assign(g->rowPtr, g->rowPtr + (size_t)g->numVertices + 1);
    } catch (const std::bad_alloc &) {
        return WTC_ERR_NOMEM;
    }
    b = wtc_big_graph{};
    b.numVertices = g->numVertices;
    b.numEdges = g->numEdges;
    b.rowPtr = rp.data();
    b.colInd = g->colInd;
    return WTC_OK;
}

int wtc_tc_arrow_gpu(const wtc_graph *g, uint32_t spread, uint32_t a, wtc_gpu_time *t, uint64_t *triangles) {
    if (!triangles) return WTC_ERR_ARG;
    std::vector<uint64_t> rp;
    wtc_big_graph b;
    const int rc = widen(g, rp, b);
    return rc ? rc : wtc_tc_arrow_gpu_big(&b, spread, a, t, triangles);
}

int wtc_tc_vertex_gpu(const wtc_graph *g, uint32_t a, int style, wtc_gpu_time *t, uint64_t *triangles) {
    std::vector<uint64_t> rp;
    wtc_big_graph b;
    const int rc = widen(g, rp, b);
    return rc ? rc : baseline_call(&b, WTC_PARALLEL_VERTEX, a, style, t, triangles);
}

int wtc_tc_vertex_gpu_big(const wtc_big_graph *g, uint32_t a, int style, wtc_gpu_time *t, uint64_t *triangles) {
    return baseline_call(g, WTC_PARALLEL_VERTEX, a, style, t, triangles);
}

int wtc_tc_edge_gpu(const wtc_graph *g, uint32_t a, int style, wtc_gpu_time *t, uint64_t *triangles) {
    std::vector<uint64_t> rp;
    wtc_big_graph b;
    const int rc = widen(g, rp, b);
    return rc ? rc : baseline_call(&b, WTC_PARALLEL_EDGE, a, style, t, triangles);
}

int wtc_tc_edge_gpu_big(const wtc_big_graph *g, uint32_t a, int style, wtc_gpu_time *t, uint64_t *triangles) {
    return baseline_call(g, WTC_PARALLEL_EDGE, a, style, t, triangles);
}

}  // extern "C"
