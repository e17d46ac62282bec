// wtc_prep.hip -- GPU preprocessing: the -p 1/2 styles of the reference's
// read_graph/preprocess (tc.cu:628-647, 697-716, 763-784, which sort with CUB)
// rebuilt on rocPRIM radix sorts plus kernels of this library, and device
// builders that leave the oriented CSR resident (including an R-MAT
// generator that draws exactly the pairs of the host generator).
//
// Every graph is a sorted set of 64-bit keys (src << bits | dst):
//   pairs -> 2 keys each (self-loops -> sentinel) -> radix sort -> unique
//   -> symmetric CSR (row bounds) -> degree per vertex -> stable radix sort of
//   (degree, id) -> rank -> keep (s,d) iff rank(d) > rank(s), relabelled
//   -> radix sort -> oriented CSR.
// The result equals the CPU builders' bit for bit (all orders are total or
// stable).
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include "wtc_dev.h"
#include "wtc_internal.h"

using namespace wtc;

namespace {

constexpr uint64_t kNoKey = ~0ull;
constexpr uint32_t kCItems = 16;
constexpr uint64_t kCChunk = (uint64_t)kBlock * kCItems;

int key_bits(uint64_t n) {
    int b = 1;
    while ((1ull << b) < n) b++;
    return b;
}

// ---------------------------------------------------------------- kernels

__global__ __launch_bounds__(kBlock) void k_pairs_to_keys(const uint32_t *__restrict__ pairs, uint64_t np, int bits,
                                                          uint64_t *__restrict__ keys) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < np; i += stride) {
        const uint64_t s = pairs[2 * i], d = pairs[2 * i + 1];
        keys[2 * i] = s == d ? kNoKey : (s << bits) | d;
        keys[2 * i + 1] = s == d ? kNoKey : (d << bits) | s;
    }
}

// The host generator's R-MAT draw (graph_host.cpp, RmatGen): pair i picks a
// quadrant per level from the counter-based stream (seed, 'RMAT', i*64+l).
__device__ __forceinline__ uint64_t mix64_d(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void k_rmat_keys(uint64_t np, uint32_t scale, uint64_t seed_mix, uint64_t ta,
                                                      uint64_t tab, uint64_t tabc, uint64_t *__restrict__ keys) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < np; i += stride) {
        uint64_t src = 0, dst = 0, r = 0;
        for (uint32_t l = 0; l < scale; l++) {
            if ((l & 1u) == 0) r = mix64_d(seed_mix + (i * 64 + l) * 0x9e3779b97f4a7c15ULL);
            const uint64_t x = (l & 1u) ? (r >> 32) : (r & 0xffffffffULL);
            const uint64_t sb = x >= tab, db = (x >= ta && x < tab) || x >= tabc;
            src = (src << 1) | sb;
            dst = (dst << 1) | db;
        }
        keys[2 * i] = src == dst ? kNoKey : (src << scale) | dst;
        keys[2 * i + 1] = src == dst ? kNoKey : (dst << scale) | src;
    }
}

struct UniqueValid {
    __device__ bool operator()(const uint64_t *in, uint64_t i) const {
        return in[i] != kNoKey && (i == 0 || in[i] != in[i - 1]);
    }
    __device__ uint64_t map(uint64_t k) const { return k; }
};

struct Orient {
    const uint32_t *rank;
    int bits;
    uint64_t mask;
    __device__ bool operator()(const uint64_t *in, uint64_t i) const {
        return rank[in[i] & mask] > rank[in[i] >> bits];
    }
    __device__ uint64_t map(uint64_t k) const {
        return ((uint64_t)rank[k >> bits] << bits) | rank[k & mask];
    }
};

template <class P>
__global__ __launch_bounds__(kBlock) void k_cmp_count(const uint64_t *__restrict__ in, uint64_t n, P p,
                                                      uint64_t *__restrict__ blk) {
    __shared__ uint64_t s_tot[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kCChunk;
    uint64_t acc = 0;
    for (uint32_t k = 0; k < kCItems; k++) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        if (i < n && p(in, i)) acc++;
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) s_tot[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) blk[blockIdx.x] = s_tot[0] + s_tot[1] + s_tot[2] + s_tot[3];
}

template <class P>
__global__ __launch_bounds__(kBlock) void k_cmp_write(const uint64_t *__restrict__ in, uint64_t n, P p,
                                                      const uint64_t *__restrict__ blk, uint64_t *__restrict__ out,
                                                      uint64_t *__restrict__ total) {
    __shared__ uint64_t s_tot[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kCChunk;
    uint64_t carry = blk[blockIdx.x];
    for (uint32_t k = 0; k < kCItems; k++) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        const bool fl = i < n && p(in, i);
        uint64_t tot;
        const uint64_t ex = block_excl_scan(fl ? 1u : 0u, s_tot, tot);
        if (fl) out[carry + ex] = p.map(in[i]);
        carry += tot;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = carry;
}

// Rows of a run of empty rows that the entry after the run writes itself;
// rows further back in a longer run are left to k_row_fill.
constexpr uint64_t kRunInline = 64;

// rowPtr[r] = first entry of row r, from sorted keys; rows without entries
// repeat the next start; rowPtr[n] = cnt.  Entry i writes the rows that end
// at its own row (the empty rows before it and its row), at most kRunInline
// of them: a long run (the isolated vertices, which the orientation ranks
// first) would otherwise be written by one lane.  rowPtr starts all ones.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_row_bounds(const uint64_t *__restrict__ keys, uint64_t cnt, int bits,
                                                       uint32_t n, OffT *__restrict__ rowPtr) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i <= cnt; i += stride) {
        const uint64_t cur = i < cnt ? (keys[i] >> bits) : n;
        const uint64_t r0 = i > 0 ? (keys[i - 1] >> bits) + 1 : 0;
        if (r0 > cur) continue;  // entry i continues the row of entry i-1
        const uint64_t lo = cur - r0 >= kRunInline ? cur + 1 - kRunInline : r0;
        for (uint64_t r = lo; r <= cur; r++) rowPtr[r] = (OffT)i;
    }
}

// The rows k_row_bounds left unset (inside long runs of empty rows, so r < n
// and r < 2^bits): the first entry whose row is >= r, by a lower-bound search
// of the sorted keys.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_row_fill(const uint64_t *__restrict__ keys, uint64_t cnt, int bits,
                                                     uint32_t n, OffT *__restrict__ rowPtr) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < n; r += stride) {
        if (rowPtr[r] != (OffT)~(OffT)0) continue;
        const uint64_t t = r << bits;
        uint64_t lo = 0, hi = cnt;
        while (lo < hi) {
            const uint64_t mid = lo + ((hi - lo) >> 1);
            if (keys[mid] < t) lo = mid + 1;
            else hi = mid;
        }
        rowPtr[r] = (OffT)lo;
    }
}

__global__ __launch_bounds__(kBlock) void k_low_bits(const uint64_t *__restrict__ keys, uint64_t cnt, uint64_t mask,
                                                     uint32_t *__restrict__ col) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < cnt; i += stride) col[i] = (uint32_t)(keys[i] & mask);
}

__global__ __launch_bounds__(kBlock) void k_degrees(const uint64_t *__restrict__ rowPtr, uint32_t n,
                                                    uint32_t *__restrict__ deg, uint32_t *__restrict__ ids) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += stride) {
        deg[v] = (uint32_t)(rowPtr[v + 1] - rowPtr[v]);
        ids[v] = (uint32_t)v;
    }
}

__global__ __launch_bounds__(kBlock) void k_ranks(const uint32_t *__restrict__ order, uint32_t n,
                                                  uint32_t *__restrict__ rank) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < n; r += stride) rank[order[r]] = (uint32_t)r;
}

// One thread per row: keys (v << bits | colInd[j]) of a symmetric CSR.
__global__ __launch_bounds__(kBlock) void k_csr_to_keys(const uint64_t *__restrict__ rowPtr,
                                                        const uint32_t *__restrict__ colInd, uint32_t n, int bits,
                                                        uint64_t *__restrict__ keys) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += stride)
        for (uint64_t j = rowPtr[v]; j < rowPtr[v + 1]; j++) keys[j] = (v << bits) | colInd[j];
}

// ------------------------------------------------------------------ host

dim3 grid_for(uint64_t n) {
    const uint64_t b = (n + kBlock - 1) / kBlock;
    return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 65536)));
}

#define PREP_CHECK(call)                                        \
    do {                                                        \
        hipError_t _e = (call);                                 \
        if (_e != hipSuccess) return hip_fail(#call, _e);       \
    } while (0)

// rowPtr[0..n] of the sorted keys[0..cnt) (rows = keys >> bits).
template <typename OffT>
hipError_t row_bounds(hipStream_t st, const uint64_t *keys, uint64_t cnt, int bits, uint32_t n, OffT *rowPtr) {
    hipError_t e = hipMemsetAsync(rowPtr, 0xFF, ((uint64_t)n + 1) * sizeof(OffT), st);
    if (e != hipSuccess) return e;
    k_row_bounds<OffT><<<grid_for(cnt + 1), dim3(kBlock), 0, st>>>(keys, cnt, bits, n, rowPtr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_row_fill<OffT><<<grid_for(n), dim3(kBlock), 0, st>>>(keys, cnt, bits, n, rowPtr);
    return hipGetLastError();
}

// Stable LSD radix sort of 64-bit keys (bits [0, end_bit)) from `in` to `out`.
int sort_keys(hipStream_t st, DevBuf &tmp, const uint64_t *in, uint64_t *out, uint64_t n, int end_bit) {
    if (n == 0) return WTC_OK;
    size_t bytes = 0;
    PREP_CHECK(rocprim::radix_sort_keys(nullptr, bytes, in, out, (size_t)n, 0u, (unsigned)end_bit, st));
    PREP_CHECK(tmp.ensure(bytes));
    PREP_CHECK(rocprim::radix_sort_keys(tmp.p, bytes, in, out, (size_t)n, 0u, (unsigned)end_bit, st));
    return WTC_OK;
}

// Ordered compaction out[...] = p.map(in[i]) for every i with p(in, i).
template <class P>
int compact(hipStream_t st, DevBuf &blk, DevBuf &total, const uint64_t *in, uint64_t n, P p, uint64_t *out,
            uint64_t *count) {
    const uint64_t nb = (n + kCChunk - 1) / kCChunk;
    if (nb == 0) {
        *count = 0;
        return WTC_OK;
    }
    if (nb > 0xffffffffULL) return WTC_ERR_RANGE;
    PREP_CHECK(blk.ensure(nb * sizeof(uint64_t)));
    PREP_CHECK(total.ensure(sizeof(uint64_t)));
    k_cmp_count<P><<<dim3((uint32_t)nb), dim3(kBlock), 0, st>>>(in, n, p, blk.as<uint64_t>());
    PREP_CHECK(hipGetLastError());
    PREP_CHECK(launch_scan_partials(blk.as<uint64_t>(), (uint32_t)nb, st));
    k_cmp_write<P><<<dim3((uint32_t)nb), dim3(kBlock), 0, st>>>(in, n, p, blk.as<uint64_t>(), out,
                                                                   total.as<uint64_t>());
    PREP_CHECK(hipGetLastError());
    PREP_CHECK(hipMemcpyAsync(count, total.p, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    PREP_CHECK(hipStreamSynchronize(st));
    return WTC_OK;
}

// raw keys (2 per pair, self-loops as sentinels) -> sorted unique keys in `out`.
// The unique keys are compacted back into raw's memory, which then becomes
// `out`; the sorted copy is handed to `spare` (when given) for the
// orientation to reuse, instead of freeing one buffer of that size and
// allocating another: at R-MAT 28 each such buffer is 69 GB, and mapping it
// costs more than the kernels that fill it.
int unique_sorted(hipStream_t st, DevBuf &raw, uint64_t nraw, int bits, DevBuf &out, uint64_t *ucount,
                  DevBuf *spare = nullptr) {
    DevBuf tmp, blk, total, sorted;
    int rc;
    PREP_CHECK(raw.ensure(std::max<uint64_t>(nraw, 1) * 8));
    PREP_CHECK(sorted.ensure(std::max<uint64_t>(nraw, 1) * 8));
    if ((rc = sort_keys(st, tmp, raw.as<uint64_t>(), sorted.as<uint64_t>(), nraw, 2 * bits))) return rc;
    PREP_CHECK(hipStreamSynchronize(st));
    tmp.release();
    // sentinel keys (all ones above 2*bits) sort last only if they compare
    // greater on the sorted bits: every valid key is < 2^(2*bits) <= sentinel
    if ((rc = compact(st, blk, total, sorted.as<uint64_t>(), nraw, UniqueValid{}, raw.as<uint64_t>(), ucount)))
        return rc;
    out = std::move(raw);
    if (spare) *spare = std::move(sorted);
    return WTC_OK;
}

// Sorted unique symmetric keys (n vertices) -> oriented CSR on the device.
// `spare`, when it holds at least 2*(U/2) keys past a 256-byte boundary,
// stores the oriented keys and their sorted copy instead of two new buffers.
template <typename OffT>
int orient_keys(hipStream_t st, const uint64_t *sym, uint64_t U, uint32_t n, int bits, DevBuf &rowPtr,
                DevBuf &colInd, DevBuf *spare = nullptr) {
    DevBuf srow, deg, ids, degs, order, rank, okeys, osorted, tmp, blk, total;
    const uint64_t mask = (1ull << bits) - 1;
    PREP_CHECK(srow.ensure(((uint64_t)n + 1) * 8));
    PREP_CHECK(row_bounds<uint64_t>(st, sym, U, bits, n, srow.as<uint64_t>()));
    PREP_CHECK(deg.ensure((uint64_t)n * 4));
    PREP_CHECK(ids.ensure((uint64_t)n * 4));
    PREP_CHECK(degs.ensure((uint64_t)n * 4));
    PREP_CHECK(order.ensure((uint64_t)n * 4));
    k_degrees<<<grid_for(n), dim3(kBlock), 0, st>>>(srow.as<uint64_t>(), n, deg.as<uint32_t>(), ids.as<uint32_t>());
    PREP_CHECK(hipGetLastError());
    // stable ascending order by degree, ties by id (tc.cu:697-716 with CUB)
    size_t bytes = 0;
    const unsigned dbits = (unsigned)key_bits((uint64_t)n + 1);
    PREP_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, deg.as<uint32_t>(), degs.as<uint32_t>(), ids.as<uint32_t>(),
                                         order.as<uint32_t>(), (size_t)n, 0u, dbits, st));
    PREP_CHECK(tmp.ensure(bytes));
    PREP_CHECK(rocprim::radix_sort_pairs(tmp.p, bytes, deg.as<uint32_t>(), degs.as<uint32_t>(), ids.as<uint32_t>(),
                                         order.as<uint32_t>(), (size_t)n, 0u, dbits, st));
    PREP_CHECK(rank.ensure((uint64_t)n * 4));
    k_ranks<<<grid_for(n), dim3(kBlock), 0, st>>>(order.as<uint32_t>(), n, rank.as<uint32_t>());
    PREP_CHECK(hipGetLastError());
    // keep v->w iff rank(w) > rank(v), relabelled (tc.cu:738-758)
    const uint64_t half = std::max<uint64_t>(U / 2, 1);
    const uint64_t second = (half * 8 + 255) & ~(uint64_t)255;  // byte offset of the sorted copy
    uint64_t *ok = nullptr, *os = nullptr;
    if (spare && spare->p && spare->cap >= second + half * 8) {
        ok = spare->as<uint64_t>();
        os = reinterpret_cast<uint64_t *>(spare->as<char>() + second);
    } else {
        PREP_CHECK(okeys.ensure(half * 8));
        PREP_CHECK(osorted.ensure(half * 8));
        ok = okeys.as<uint64_t>();
        os = osorted.as<uint64_t>();
    }
    uint64_t m = 0;
    int rc = compact(st, blk, total, sym, U, Orient{rank.as<uint32_t>(), bits, mask}, ok, &m);
    if (rc) return rc;
    if (m * 2 != U) return WTC_ERR_CSR;
    if ((rc = sort_keys(st, tmp, ok, os, m, 2 * bits))) return rc;
    PREP_CHECK(rowPtr.ensure(((uint64_t)n + 1) * sizeof(OffT)));
    PREP_CHECK(colInd.ensure(std::max<uint64_t>(m, 1) * 4));
    PREP_CHECK(row_bounds<OffT>(st, os, m, bits, n, rowPtr.as<OffT>()));
    k_low_bits<<<grid_for(m), dim3(kBlock), 0, st>>>(os, m, mask, colInd.as<uint32_t>());
    PREP_CHECK(hipGetLastError());
    PREP_CHECK(hipStreamSynchronize(st));
    return WTC_OK;
}

int download_csr(hipStream_t st, const DevBuf &rowPtr, int ob, const DevBuf &colInd, uint32_t n, uint64_t m,
                 wtc_big_graph **out) {
    wtc_big_graph *g = static_cast<wtc_big_graph *>(std::calloc(1, sizeof(wtc_big_graph)));
    uint64_t *rp = alloc_n<uint64_t>((uint64_t)n + 1);
    uint32_t *ci = alloc_n<uint32_t>(m);
    if (!g || !rp || !ci) {
        std::free(g);
        std::free(rp);
        std::free(ci);
        return WTC_ERR_NOMEM;
    }
    hipError_t e;
    if (ob == 8) {
        e = hipMemcpyAsync(rp, rowPtr.p, ((uint64_t)n + 1) * 8, hipMemcpyDeviceToHost, st);
    } else {
        std::vector<uint32_t> r32((uint64_t)n + 1);
        e = hipMemcpyAsync(r32.data(), rowPtr.p, ((uint64_t)n + 1) * 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        for (uint64_t v = 0; v <= n && e == hipSuccess; v++) rp[v] = r32[v];
    }
    if (e == hipSuccess && m) e = hipMemcpyAsync(ci, colInd.p, m * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        std::free(g);
        std::free(rp);
        std::free(ci);
        return hip_fail("download csr", e);
    }
    g->numVertices = n;
    g->numEdges = m;
    g->rowPtr = rp;
    g->colInd = ci;
    *out = g;
    return WTC_OK;
}

// Oriented device CSR from sorted unique symmetric keys, into a new dgraph.
int make_dgraph(wtc_ctx *c, const uint64_t *sym, uint64_t U, uint32_t n, int bits, uint32_t ob, wtc_dgraph **out,
                DevBuf *spare = nullptr) {
    wtc_dgraph *dg = new (std::nothrow) wtc_dgraph();
    if (!dg) return WTC_ERR_NOMEM;
    dg->ctx = c;
    dg->n = n;
    dg->m = U / 2;
    if (ob == 0) ob = dg->m < 0xffffffffULL ? 4 : 8;
    if (ob == 4 && dg->m >= 0xffffffffULL) {
        delete dg;
        return WTC_ERR_RANGE;
    }
    dg->ob = (int)ob;
    const int rc = ob == 8 ? orient_keys<uint64_t>(c->slot.stream, sym, U, n, bits, dg->rowPtr, dg->colInd, spare)
                           : orient_keys<uint32_t>(c->slot.stream, sym, U, n, bits, dg->rowPtr, dg->colInd, spare);
    if (rc) {
        wtc_dgraph_free(dg);
        return rc;
    }
    *out = dg;
    return WTC_OK;
}

// n of a graph from its sorted unique symmetric keys: largest id + 1 (the
// last key holds the largest source), 1 when there is no edge.
int n_from_keys(hipStream_t st, const DevBuf &keys, uint64_t U, int bits, uint32_t *n) {
    if (U == 0) {
        *n = 1;
        return WTC_OK;
    }
    uint64_t last = 0;
    PREP_CHECK(hipMemcpyAsync(&last, keys.as<uint64_t>() + (U - 1), 8, hipMemcpyDeviceToHost, st));
    PREP_CHECK(hipStreamSynchronize(st));
    *n = (uint32_t)((last >> bits) + 1);
    return WTC_OK;
}

// Symmetric keys of host pairs (ids < n_cap) on the context's stream.
int keys_from_pairs(wtc_ctx *c, const uint32_t *pairs, uint64_t np, int bits, DevBuf &sym, uint64_t *U,
                    DevBuf *spare = nullptr) {
    hipStream_t st = c->slot.stream;
    DevBuf dpairs, raw;
    PREP_CHECK(dpairs.ensure(std::max<uint64_t>(np, 1) * 8));
    PREP_CHECK(raw.ensure(std::max<uint64_t>(2 * np, 1) * 8));
    if (np) PREP_CHECK(hipMemcpyAsync(dpairs.p, pairs, np * 8, hipMemcpyHostToDevice, st));
    k_pairs_to_keys<<<grid_for(np), dim3(kBlock), 0, st>>>(dpairs.as<uint32_t>(), np, bits, raw.as<uint64_t>());
    PREP_CHECK(hipGetLastError());
    PREP_CHECK(hipStreamSynchronize(st));
    dpairs.release();
    return unique_sorted(st, raw, 2 * np, bits, sym, U, spare);
}

}  // namespace

namespace wtc {

int gpu_symmetric_from_pairs(wtc_ctx *c, const uint32_t *pairs, uint64_t np, uint32_t n, wtc_big_graph **out) {
    WTC_HIPC(hipSetDevice(c->device));
    const int bits = key_bits(n);
    DevBuf sym, srow, scol;
    uint64_t U = 0;
    int rc = keys_from_pairs(c, pairs, np, bits, sym, &U);
    if (rc) return rc;
    uint32_t nn = 1;
    if ((rc = n_from_keys(c->slot.stream, sym, U, bits, &nn))) return rc;
    hipStream_t st = c->slot.stream;
    PREP_CHECK(srow.ensure(((uint64_t)nn + 1) * 8));
    PREP_CHECK(scol.ensure(std::max<uint64_t>(U, 1) * 4));
    PREP_CHECK(row_bounds<uint64_t>(st, sym.as<uint64_t>(), U, bits, nn, srow.as<uint64_t>()));
    k_low_bits<<<grid_for(U), dim3(kBlock), 0, st>>>(sym.as<uint64_t>(), U, (1ull << bits) - 1,
                                                       scol.as<uint32_t>());
    PREP_CHECK(hipGetLastError());
    return download_csr(st, srow, 8, scol, nn, U, out);
}

int gpu_orient(wtc_ctx *c, const wtc_big_graph *g, wtc_big_graph **out) {
    WTC_HIPC(hipSetDevice(c->device));
    hipStream_t st = c->slot.stream;
    const uint32_t n = g->numVertices;
    const uint64_t U = g->numEdges;
    const int bits = key_bits(n);
    DevBuf drow, dcol, sym, orow, ocol;
    PREP_CHECK(drow.ensure(((uint64_t)n + 1) * 8));
    PREP_CHECK(dcol.ensure(std::max<uint64_t>(U, 1) * 4));
    PREP_CHECK(sym.ensure(std::max<uint64_t>(U, 1) * 8));
    PREP_CHECK(hipMemcpyAsync(drow.p, g->rowPtr, ((uint64_t)n + 1) * 8, hipMemcpyHostToDevice, st));
    if (U) PREP_CHECK(hipMemcpyAsync(dcol.p, g->colInd, U * 4, hipMemcpyHostToDevice, st));
    k_csr_to_keys<<<grid_for(n), dim3(kBlock), 0, st>>>(drow.as<uint64_t>(), dcol.as<uint32_t>(), n, bits,
                                                        sym.as<uint64_t>());
    PREP_CHECK(hipGetLastError());
    int rc = orient_keys<uint64_t>(st, sym.as<uint64_t>(), U, n, bits, orow, ocol);
    if (rc) return rc;
    return download_csr(st, orow, 8, ocol, n, U / 2, out);
}

}  // namespace wtc

// Entry points of graph_host.cpp's -p 1/2 paths (default context).
int wtc_gpu_symmetric_default(const uint32_t *pairs, uint64_t np, uint32_t n, wtc_big_graph **out) {
    std::lock_guard<std::mutex> lk(default_ctx_mutex());
    wtc_ctx *c;
    int rc = default_context(&c);
    if (rc) return rc;
    return gpu_symmetric_from_pairs(c, pairs, np, n, out);
}

int wtc_gpu_orient_default(const wtc_big_graph *g, wtc_big_graph **out) {
    std::lock_guard<std::mutex> lk(default_ctx_mutex());
    wtc_ctx *c;
    int rc = default_context(&c);
    if (rc) return rc;
    return gpu_orient(c, g, out);
}

extern "C" {

int wtc_dgraph_from_pairs(wtc_ctx *c, const uint32_t *pairs, uint64_t npairs, uint32_t offset_bytes,
                          wtc_dgraph **out) {
    if (!c || !out || (npairs && !pairs) || (offset_bytes != 0 && offset_bytes != 4 && offset_bytes != 8))
        return WTC_ERR_ARG;
    *out = nullptr;
    uint32_t mx = 0;
    for (uint64_t i = 0; i < npairs; i++)
        if (pairs[2 * i] != pairs[2 * i + 1]) mx = std::max(mx, std::max(pairs[2 * i], pairs[2 * i + 1]));
    if ((uint64_t)mx + 1 > 0xffffffffULL) return WTC_ERR_FORMAT;
    WTC_HIPC(hipSetDevice(c->device));
    const int bits = key_bits((uint64_t)mx + 1);
    DevBuf sym, spare;
    uint64_t U = 0;
    int rc = keys_from_pairs(c, pairs, npairs, bits, sym, &U, &spare);
    if (rc) return rc;
    uint32_t n = 1;
    if ((rc = n_from_keys(c->slot.stream, sym, U, bits, &n))) return rc;
    return make_dgraph(c, sym.as<uint64_t>(), U, n, bits, offset_bytes, out, &spare);
}

int wtc_dgraph_gen_rmat(wtc_ctx *c, uint32_t scale, uint32_t ef, double a, double b, double cc, uint64_t seed,
                        uint32_t offset_bytes, wtc_dgraph **out) {
    if (!c || !out || scale == 0 || scale > 31 || ef == 0 || a < 0 || b < 0 || cc < 0 || a + b + cc > 1.0 ||
        (offset_bytes != 0 && offset_bytes != 4 && offset_bytes != 8))
        return WTC_ERR_ARG;
    *out = nullptr;
    WTC_HIPC(hipSetDevice(c->device));
    hipStream_t st = c->slot.stream;
    const double two32 = 4294967296.0;
    const uint64_t np = (uint64_t)ef << scale;
    const int bits = (int)scale;
    DevBuf raw, sym;
    PREP_CHECK(raw.ensure(2 * np * 8));
    const uint64_t seed_mix = mix64(seed * 0x9e3779b97f4a7c15ULL + 0x524d4154ULL);
    k_rmat_keys<<<grid_for(np), dim3(kBlock), 0, st>>>(np, scale, seed_mix, (uint64_t)(a * two32),
                                                       (uint64_t)((a + b) * two32), (uint64_t)((a + b + cc) * two32),
                                                       raw.as<uint64_t>());
    PREP_CHECK(hipGetLastError());
    uint64_t U = 0;
    DevBuf spare;
    int rc = unique_sorted(st, raw, 2 * np, bits, sym, &U, &spare);
    if (rc) return rc;
    uint32_t n = 1;
    if ((rc = n_from_keys(st, sym, U, bits, &n))) return rc;
    return make_dgraph(c, sym.as<uint64_t>(), U, n, bits, offset_bytes, out, &spare);
}

int wtc_dgraph_download(const wtc_dgraph *dg, wtc_big_graph **out) {
    if (!dg || !out) return WTC_ERR_ARG;
    *out = nullptr;
    WTC_HIPC(hipSetDevice(dg->ctx->device));
    return download_csr(dg->ctx->slot.stream, dg->rowPtr, dg->ob, dg->colInd, dg->n, dg->m, out);
}

}  // extern "C"
