// wtc_kernels.hip -- HIP kernels of the wedge-parallel triangle counter for
// CDNA4 (gfx950, wave64).
//
// Device pipeline for a graph set (GraphDesc array; one graph or many):
//   k_wedge_partials / k_scan_partials / k_wedge_scan
//       per graph, wedgeSum[v] = sum_{x<v} C(d+(x),2) (the host prefix of
//       tc.cu:348-361, as one reduce-then-scan over all graphs' vertices)
//   k_setup       per-graph wedge ranges, tile bases, the 12-level search
//                 tree over wedgeSum (tc.cu:306-313) and zeroed counts
//   k_bitmap_clear / k_bitmap_build
//                 the strict-upper-triangular bitmap over each graph's top-a
//                 ranks (tc.cu:329-346), built with atomicOr
//   k_count       the wedge-parallel count (tc.cu:173-293): persistent,
//                 XCD-partitioned grid; every wave owns tiles of 64*s wedges
//                 and keeps its own LDS staging, so waves never wait on each
//                 other inside a tile.
//   k_count_small the same count for a set of small graphs, one block per
//                 slice of a graph's wedges with the graph's CSR in LDS.
#include "wtc_dev.h"
#include "wtc_kernels.h"
#include "wtc_launch.h"

namespace wtc {

// Most tiles a wave of k_count takes per grab from its XCD's work counter.
// Every grab is one 64-bit atomic on the counter its XCD's waves share, and
// those atomics are serialised at the L2: with 8 tiles per grab they alone
// set a floor of ~80 ms on RMAT-24 (54M tiles); with 32, ~20 ms.  Launches
// with few tiles take smaller grabs (about two per wave), so that a small
// graph still spreads over many waves.
constexpr uint64_t kTilesPerGrabMax = 32;

// Rightmost graph g with desc[g].*field <= key (wave-uniform, scalar loads).
__device__ __forceinline__ uint32_t graph_of(const GraphDesc *__restrict__ desc, uint32_t G, uint64_t key,
                                            uint64_t GraphDesc::*field) {
    uint32_t lo = 0, hi = G - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (desc[mid].*field <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------ wedge prefix

// Block b scans vertices [lb*kScanChunk, (lb+1)*kScanChunk) of graph g, where
// b = desc[g].blk_base + lb.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_wedge_partials(const GraphDesc *__restrict__ desc, uint32_t G,
                                                           uint64_t *__restrict__ partials) {
    __shared__ uint64_t s_tot[kWavesPerBlock];
    const uint32_t g = G > 1 ? graph_of(desc, G, blockIdx.x, &GraphDesc::blk_base) : 0u;
    const GraphDesc &D = desc[g];
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(D.rowPtr);
    const uint32_t n = D.n;
    const uint64_t base = (blockIdx.x - D.blk_base) * (uint64_t)kScanChunk;
    uint64_t acc = 0;
#pragma unroll 4
    for (int k = 0; k < kScanItems; k++) {
        const uint64_t v = base + (uint64_t)k * kBlock + threadIdx.x;
        if (v < n) acc += binom2_d((uint64_t)(rowPtr[v + 1] - rowPtr[v]));
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) s_tot[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = s_tot[0] + s_tot[1] + s_tot[2] + s_tot[3];
}

__global__ __launch_bounds__(kBlock) void k_scan_partials(uint64_t *__restrict__ partials, uint32_t nb) {
    __shared__ uint64_t s_tot[kWavesPerBlock];
    uint64_t carry = 0;
    for (uint32_t base = 0; base < nb; base += kBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint64_t x = i < nb ? partials[i] : 0;
        uint64_t tot;
        const uint64_t ex = block_excl_scan(x, s_tot, tot);
        if (i < nb) partials[i] = carry + ex;
        carry += tot;
    }
}

template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_wedge_scan(const GraphDesc *__restrict__ desc, uint32_t G,
                                                       const uint64_t *__restrict__ partials) {
    __shared__ uint64_t s_tot[kWavesPerBlock];
    const uint32_t g = G > 1 ? graph_of(desc, G, blockIdx.x, &GraphDesc::blk_base) : 0u;
    const GraphDesc &D = desc[g];
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(D.rowPtr);
    uint64_t *__restrict__ ws = D.ws;
    const uint32_t n = D.n;
    const uint64_t lb = blockIdx.x - D.blk_base;
    uint64_t carry = partials[blockIdx.x] - partials[D.blk_base];  // graph-local prefix
    const uint64_t base = lb * kScanChunk;
    for (int k = 0; k < kScanItems; k++) {
        if (base + (uint64_t)k * kBlock >= n) break;  // block-uniform: past the graph's last vertex
        const uint64_t v = base + (uint64_t)k * kBlock + threadIdx.x;
        const uint64_t x = v < n ? binom2_d((uint64_t)(rowPtr[v + 1] - rowPtr[v])) : 0;
        uint64_t tot;
        const uint64_t ex = block_excl_scan(x, s_tot, tot);
        if (v < n) ws[v] = carry + ex;
        carry += tot;
    }
    if (lb == (n - 1) / kScanChunk && threadIdx.x == 0) ws[n] = carry;  // the graph's last block
}

// ------------------------------------------------------------------- setup

// One block.  Per graph: wedge range, tile count (64*s wedges per tile) and
// exclusive tile base; meta totals; counts zeroed; search tree for a single
// graph (node i = wedgeSum at the midpoint of its range, tc.cu:306-313).
__global__ __launch_bounds__(kBlock) void k_setup(GraphDesc *__restrict__ desc, uint32_t G,
                                                  uint64_t *__restrict__ tree, uint64_t *__restrict__ meta,
                                                  unsigned long long *__restrict__ counts) {
    __shared__ uint64_t s_tot[kWavesPerBlock];
    uint64_t carry = 0, wsum = 0;
    for (uint32_t base = 0; base < G; base += kBlock) {
        const uint32_t g = base + threadIdx.x;
        uint64_t tiles = 0, wc = 0;
        if (g < G) {
            GraphDesc &D = desc[g];
            const uint64_t W = D.ws[D.n];
            const uint64_t b = D.w_begin < W ? D.w_begin : W;
            const uint64_t e = D.w_end < W ? D.w_end : W;
            wc = e > b ? e - b : 0;
            D.wbase = b;
            D.wcount = wc;
            const uint64_t per = (uint64_t)kWave * D.spread;
            // small graphs: k_count_small; invalid ones: not counted
            tiles = D.mode == kModeTiles ? (wc + per - 1) / per : 0;
            counts[g] = D.mode == kModeInvalid ? kCountInvalid : 0ull;
        }
        uint64_t tot, wtot;
        const uint64_t ex = block_excl_scan(tiles, s_tot, tot);
        if (g < G) desc[g].tile_base = carry + ex;
        carry += tot;
        block_excl_scan(wc, s_tot, wtot);
        wsum += wtot;
    }
    if (threadIdx.x == 0) {
        meta[kMetaTiles] = carry;
        meta[kMetaWedges] = wsum;
    }
    if (threadIdx.x < 8) meta[kMetaNext + threadIdx.x] = 0;  // per-XCD tile counters of k_count
    if (tree != nullptr) {  // single-graph sets only
        const uint32_t n = desc[0].n;
        const uint64_t *ws = desc[0].ws;
        for (uint32_t i = threadIdx.x; i < kTreeSize; i += kBlock) {
            const uint32_t h = i + 1;
            const int L = 31 - __clz(h);
            uint64_t s = 0, e = n;
            for (int bit = L - 1; bit >= 0; bit--) {
                const uint64_t mid = (s + e) >> 1;
                if ((h >> bit) & 1u) s = mid + 1;
                else e = mid;
            }
            tree[i] = ws[(s + e) >> 1];
        }
    }
}

// ------------------------------------------------------------------ bitmap

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t *__restrict__ p, uint64_t n, uint32_t v) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Zeroes the bitmaps of every graph of the set: word i of the set's clear
// list belongs to the rightmost graph whose bm_clear_base <= i.
__global__ __launch_bounds__(kBlock) void k_bitmap_clear(const GraphDesc *__restrict__ desc, uint32_t G,
                                                         uint64_t total) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += stride) {
        uint32_t lo = 0, hi = G - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (desc[mid].bm_clear_base <= i) lo = mid;
            else hi = mid - 1;
        }
        desc[lo].bitmap[i - desc[lo].bm_clear_base] = 0u;
    }
}

// One wave per bitmap row (rows v in [n - min(n,a), n) of every graph): the
// device form of the host loop tc.cu:329-346, in the square layout of
// bm_word (a row owns its words alone).
template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_bitmap_build(const GraphDesc *__restrict__ desc, uint32_t G,
                                                         uint64_t rows) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t gw = (uint64_t)blockIdx.x * kWavesPerBlock + uni32(threadIdx.x >> 6);
    const uint64_t nw = (uint64_t)gridDim.x * kWavesPerBlock;
    for (uint64_t r = gw; r < rows; r += nw) {
        const uint32_t g = G > 1 ? graph_of(desc, G, r, &GraphDesc::row_base) : 0u;
        const GraphDesc &D = desc[g];
        if (D.mode == kModeInvalid) continue;  // wave-uniform: its rows are never read
        const OffT *__restrict__ rowPtr = static_cast<const OffT *>(D.rowPtr);
        const uint32_t *__restrict__ colInd = D.colInd;
        uint32_t *__restrict__ bm = D.bitmap;
        const uint32_t v = D.n - D.bm_rows + (uint32_t)(r - D.row_base);
        const uint32_t a32 = D.a >> 5;
        const OffT b = rowPtr[v], e = rowPtr[v + 1];
        for (OffT j0 = b; j0 < e; j0 += kWave) {  // wave-uniform
            const OffT j = j0 + lane;
            uint64_t word = ~0ull;  // no lane's word: inactive lanes never match
            uint32_t acc = 0;
            if (j < e) {
                const uint32_t u = colInd[j];
                word = bm_word(v, u, D.thr, a32);
                acc = 1u << ((u - D.thr) & 31u);
            }
            // ids ascend along the lanes, so lanes sharing a word are
            // consecutive: OR each run into its last lane (segmented scan)
#pragma unroll
            for (int dd = 1; dd < kWave; dd <<= 1) {
                const uint32_t ov = (uint32_t)__shfl_up((int)acc, dd, kWave);
                const uint64_t ow = shfl_up_u64(word, dd);
                if ((int)lane >= dd && ow == word) acc |= ov;
            }
            const uint64_t next_w = shfl_u64(word, lane + 1 < (uint32_t)kWave ? lane + 1 : lane);
            const uint64_t first_w = shfl_u64(word, 0);
            const bool run_end = j < e && (lane == kWave - 1 || next_w != word);
            if (run_end) {
                // a run reaching lane 63 may continue in the next 64-edge
                // chunk and one holding lane 0 may have begun in the previous
                // one: atomic; every other word belongs to this chunk alone
                if (lane == kWave - 1 || word == first_w) atomicOr(&bm[word], acc);
                else bm[word] = acc;
            }
        }
    }
}

// ------------------------------------------------------------------- count

// Wave-uniform search for the tile's first vertex over [0, n) of a single
// graph: 12 levels from the compact search tree, then wedgeSum
// (tc.cu:139-170 semantics).
template <class TP, class P>
__device__ __forceinline__ uint32_t tree_search(TP tree, P ws, uint32_t n, uint64_t t) {
    uint32_t gs = 0, ge = n, ci = 0;
#pragma unroll
    for (int l = 0; l < kTreeLevels; l++) {
        const uint64_t mv = tree[ci];
        const uint32_t gm = (uint32_t)(((uint64_t)gs + ge) >> 1);
        ci = 2 * ci + 1;
        if (mv <= t) {
            ci += 1;
            gs = gm + 1;
        } else {
            ge = gm;
        }
    }
    gs = gs > 0 ? gs - 1 : 0;
    // rightmost in [gs, ge) with ws <= t (ws[gs] <= t holds)
    return ge > gs ? rightmost_le(ws, gs, ge - 1, t) : gs;
}

// Wave-uniform galloping search from a known lower bound.
template <class P>
__device__ __forceinline__ uint32_t gallop_le(P ws, uint32_t lo, uint32_t vmax, uint64_t t) {
    uint64_t step = 1;
    for (;;) {
        const uint64_t p = (uint64_t)lo + step;
        if (p > vmax || ws[p] > t) break;
        lo = (uint32_t)p;
        step <<= 1;
    }
    const uint64_t hi = (uint64_t)lo + step - 1;
    return rightmost_le(ws, lo, hi < vmax ? (uint32_t)hi : vmax, t);
}

// Local wedge index iv of a vertex with out-degree d (d >= 2, iv < C(d,2))
// -> positions (w_i, u_i), 0 <= w_i < u_i < d, in row-major order over the
// strict upper triangle (the mapping of tc.cu:209-212).  Integer-exact for
// every d < 2^32: a double-precision estimate of the row, then exact fix-up
// against binomial counts.
__device__ __forceinline__ void decode_wedge(uint32_t d, uint64_t iv, uint32_t &w_i, uint32_t &u_i) {
    const uint64_t C = binom2_d(d);
    const uint64_t kl = C - iv;  // wedges from iv to the end, >= 1
    // q = d-1-w_i is the largest q with C(q,2) < kl
    uint64_t q = (uint64_t)(__builtin_sqrt(2.0 * (double)kl) + 0.5);
    q = q < 1 ? 1 : (q > (uint64_t)d - 1 ? (uint64_t)d - 1 : q);
    while (binom2_d(q) >= kl) q--;
    while (binom2_d(q + 1) < kl) q++;
    w_i = d - 1 - (uint32_t)q;
    u_i = w_i + 1 + (uint32_t)(iv - (C - binom2_d(q + 1)));
}

__global__ __launch_bounds__(kBlock) void k_debug_decode(const uint32_t *__restrict__ d,
                                                         const uint64_t *__restrict__ iv, uint64_t k,
                                                         uint32_t *__restrict__ w_out,
                                                         uint32_t *__restrict__ u_out) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= k) return;
    uint32_t w, u;
    decode_wedge(d[i], iv[i], w, u);
    w_out[i] = w;
    u_out[i] = u;
}

// Four ids loaded by one 16-byte vector load, at any element (4-byte)
// alignment, which the global memory instructions accept for multi-dword
// loads.
typedef uint32_t u32x4_e __attribute__((ext_vector_type(4), aligned(4)));

// kWays: independent probes each lane advances together.  The block holds
// blockDim.x / 64 waves (1..kCountWavesMax), each with its own LDS staging.
template <typename OffT, int kWays>
__global__ __launch_bounds__(kWave * kCountWavesMax) void k_count(const GraphDesc *__restrict__ desc, uint32_t G,
                                                                  const uint64_t *__restrict__ tree,
                                                                  uint64_t *__restrict__ meta,
                                                                  unsigned long long *__restrict__ counts,
                                                                  uint32_t pad_max, uint32_t chunk, uint32_t work) {
    extern __shared__ uint2 lds_wu[];
    __shared__ uint32_t s_g[kCountWavesMax];
    __shared__ uint64_t s_c[kCountWavesMax];

    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = uni32(threadIdx.x >> 6);
    const uint32_t wpb = blockDim.x >> 6;  // waves of this block
    // this wave's staging: (w, u) of a staged wedge as one 8-byte entry,
    // lane l's k-th wedge at l * pad + k
    uint2 *const lwu = lds_wu + wid * (kWave * pad_max);

    // XCD interleave: blocks go round-robin to the 8 XCDs, so blocks with the
    // same blockIdx % 8 share an L2.  The tiles are cut into super-chunks of
    // 2^lg grabs (tpg consecutive tiles each), dealt to the XCDs in
    // turn: XCD x owns super-chunks x, x+8, x+16, ..., so every XCD gets an
    // even mix of cheap and expensive regions of the wedge space (the cost per
    // wedge varies several-fold along it), while each super-chunk stays
    // contiguous for its L2.  Inside an XCD, waves take grabs from the XCD's
    // own counter (dynamic balance); consecutive tiles of a grab let a wave
    // start its vertex window where its previous tile ended.
    const uint64_t ntiles = meta[kMetaTiles];
    // grab size: about two grabs per wave, at most kTilesPerGrabMax tiles;
    // super-chunks of 2^lg grabs, the largest power of two (at most 64) that
    // still gives every XCD at least 64 super-chunks
    const uint64_t waves = (uint64_t)gridDim.x * wpb;
    uint64_t tpg = ntiles / (2 * waves);
    const uint64_t gmax = (work & 0xffffu) ? (work & 0xffffu) : kTilesPerGrabMax;
    tpg = tpg < 1 ? 1 : tpg > gmax ? gmax : tpg;
    const uint64_t ngrabs = (ntiles + tpg - 1) / tpg;
    // dealing to the XCDs (work bit 16) or one stream for the whole grid
    const bool deal = (work >> 16) & 1u;
    const uint32_t streams = deal ? 8u : 1u;
    uint32_t lg = 0;
    while (deal && lg < 6 && (ngrabs >> (lg + 1)) >= 8ull * 64) lg++;
    const uint32_t xcd = deal ? blockIdx.x & 7u : 0u;
    unsigned long long *const ctr = reinterpret_cast<unsigned long long *>(meta + kMetaNext + xcd);

    uint32_t cur_g = kSentinel;
    uint64_t cur_cnt = 0;

    for (;;) {
        unsigned long long grab = 0;
        if (lane == 0) grab = atomicAdd(ctr, 1ull);
        const uint64_t gi = uni64(grab);
        const uint64_t chunk_id = xcd + streams * (gi >> lg);  // this XCD's super-chunk
        const uint64_t tb = ((chunk_id << lg) + (gi & ((1ull << lg) - 1))) * tpg;
        if (tb >= ntiles) break;  // tb grows with gi, so every later grab is past the end too
        const uint64_t te = tb + tpg < ntiles ? tb + tpg : ntiles;

        uint32_t g = kSentinel, v_prev = 0;
        uint64_t g_tile_end = 0;
        bool have_prev = false;
        // per-graph state, wave-uniform
        gptr<OffT> rowPtr = nullptr;
        gptr<uint32_t> colInd = nullptr;
        gptr<uint64_t> ws = nullptr;
        gptr<uint32_t> bitmap = nullptr;
        uint32_t s = 1, sc = 1, pad = 1, gn = 1, thr = 0;
        uint32_t a32 = 0, bm_rows = 0;
        bool bm_narrow = true;
        uint64_t wbase = 0, wend = 0, tile_base = 0;

        for (uint64_t t = tb; t < te; t++) {
            if (g == kSentinel || t >= g_tile_end) {
                g = G > 1 ? graph_of(desc, G, t, &GraphDesc::tile_base) : 0u;
                const GraphDesc &D = desc[g];
                rowPtr = to_global(static_cast<const OffT *>(D.rowPtr));
                colInd = to_global(D.colInd);
                ws = to_global(static_cast<const uint64_t *>(D.ws));
                bitmap = to_global(static_cast<const uint32_t *>(D.bitmap));
                s = D.spread;
                sc = s < chunk ? s : chunk;  // wedges per lane staged at once
                pad = sc | 1u;
                gn = D.n;
                thr = D.thr;
                a32 = D.a >> 5;
                bm_rows = D.bm_rows;
                bm_narrow = D.bm_words < (1ull << 32);
                wbase = D.wbase;
                wend = D.wbase + D.wcount;
                tile_base = D.tile_base;
                const uint64_t per = (uint64_t)kWave * s;
                g_tile_end = tile_base + (D.wcount + per - 1) / per;
                have_prev = false;
            }
            const uint32_t vmax = gn - 1;
            const uint64_t tile_start = wbase + (t - tile_base) * ((uint64_t)kWave * s);
            const uint64_t tile_end =
                tile_start + (uint64_t)kWave * s < wend ? tile_start + (uint64_t)kWave * s : wend;

            // Vertex range [v_lo, v_hi] of the tile and each lane's first
            // vertex from one coalesced window: wedgeSum and rowPtr of the 64
            // vertices from `base` (the previous tile's last vertex, or a
            // search for the first tile of a grab), lane l holding base + l.
            // wedgeSum is non-decreasing, so ballots of `<= key` are prefixes
            // whose top bit gives the range; a lane finds its vertex by a
            // 6-step search over the window's lanes and its row bounds by
            // shuffles.  A tile that reaches past the window falls back to
            // searches in global memory.
            uint32_t base = have_prev ? v_prev
                            : (tree != nullptr) ? tree_search(tree, ws, gn, tile_start)
                                                : rightmost_le(ws, 0u, vmax, tile_start);
            uint64_t wwin, m_lo;
            OffT rwin;
            for (;;) {  // ws[base] <= tile_start; at most one skip ahead
                const uint32_t x = base + lane;
                wwin = x <= vmax ? ws[x] : ~0ull;
                rwin = x <= gn ? rowPtr[x] : (OffT)0;
                m_lo = __ballot(wwin <= tile_start);
                if (~m_lo != 0) break;
                base = gallop_le(ws, base + kWave - 1, vmax, tile_start);  // all 64 at or below tile_start
            }
            const uint32_t lo_l = 63u - (uint32_t)__clzll(m_lo);
            const uint64_t m_hi = __ballot(wwin <= tile_end - 1);
            const bool hi_in = ~m_hi != 0;  // wave-uniform
            const uint32_t v_hi = hi_in ? base + (63u - (uint32_t)__clzll(m_hi))
                                        : gallop_le(ws, base + kWave - 1, vmax, tile_end - 1);
            v_prev = v_hi;
            have_prev = true;

            // lane state: wedges [first, lend); the first one is located in
            // the window (or past it) and decoded exactly, the rest by stepping
            const uint64_t first = tile_start + (uint64_t)lane * s;
            const uint64_t lend = first + s < tile_end ? first + s : tile_end;
            // wedges of this lane not yet staged (<= s, so 32 bits); the 64-bit
            // index of the next one is lend - left
            uint32_t left = first < lend ? (uint32_t)(lend - first) : 0u;
            uint32_t pos = lo_l;
            uint64_t iv = 0;  // local index of the lane's first wedge in its vertex
            if (s < (1u << 25)) {  // wave-uniform: every lane offset lane*s fits 31 bits
                // The window's prefixes relative to the tile start, saturated
                // to 32 bits: 0 for the vertices starting at or before it,
                // ~0 for those 2^32 or more wedges on (and past the graph).
                // Compared with the lane's offset lane*s < 2^31 they order
                // exactly as the 64-bit prefixes do against its first wedge.
                const uint64_t dw = wwin - tile_start;
                const uint32_t rel = wwin <= tile_start ? 0u : dw > 0xffffffffull ? 0xffffffffu : (uint32_t)dw;
                const uint32_t foff = lane * s;
                // rightmost window index in [lo_l, 63] with prefix <= first
                // (binary lifting; every lane takes part in the shuffles)
#pragma unroll
                for (uint32_t step = 32; step > 0; step >>= 1) {
                    const uint32_t cand = pos + step;
                    const uint32_t rc = shfl_off(rel, cand < (uint32_t)kWave ? cand : kWave - 1);
                    pos = cand < (uint32_t)kWave && rc <= foff ? cand : pos;
                }
                const uint32_t rel_v = shfl_off(rel, pos);
                // the tile's first vertex may start before the tile: its exact
                // prefix is the wave-uniform window value at lo_l
                const uint64_t ws_lo = readlane_u64(wwin, lo_l);
                iv = pos == lo_l ? first - ws_lo : (uint64_t)(foff - rel_v);
            } else {
#pragma unroll
                for (uint32_t step = 32; step > 0; step >>= 1) {
                    const uint32_t cand = pos + step;
                    const uint64_t wc = shfl_u64(wwin, cand < (uint32_t)kWave ? cand : kWave - 1);
                    pos = cand < (uint32_t)kWave && wc <= first ? cand : pos;
                }
                iv = first - shfl_u64(wwin, pos);
            }
            const OffT rb = shfl_off(rwin, pos);
            const OffT re = shfl_off(rwin, pos + 1 < (uint32_t)kWave ? pos + 1 : kWave - 1);
            uint32_t v = 0, d = 0, w_i = 0, u_i = 0;
            OffT vb = 0;
            if (left) {
                if (pos == kWave - 1 && !hi_in) {  // the lane's vertex lies past the window
                    v = rightmost_le(ws, base + kWave - 1, v_hi, first);
                    vb = rowPtr[v];
                    d = (uint32_t)(rowPtr[v + 1] - vb);
                    decode_wedge(d, first - ws[v], w_i, u_i);
                } else {
                    v = base + pos;
                    vb = rb;
                    d = (uint32_t)((pos + 1 < (uint32_t)kWave ? re : rowPtr[v + 1]) - vb);
                    decode_wedge(d, iv, w_i, u_i);
                }
            }
            // hits of the wave in this tile: popcount(ballot(match)) as each
            // probe step resolves (a probe matches at most once); the ballot
            // is the compare's own lane mask, so counting costs scalar
            // instructions only
            uint64_t tile_cnt = 0;
            // colInd position and value of the last w loaded: consecutive
            // wedges of a lane mostly share w, whose id is then not reloaded
            OffT last_wpos = ~(OffT)0;
            uint32_t last_wv = 0;

            for (uint64_t c = 0; c < s; c += sc) {  // 64-bit: s may be close to 2^32
                const uint32_t kn = (s - c) < sc ? (uint32_t)(s - c) : sc;
                // phase 1: this lane's next kn wedges into LDS, positions first
                // and then the colInd loads of a group together
                for (uint32_t k0 = 0; k0 < kn; k0 += 4) {
                    // Fast path, taken by the whole wave when every lane that
                    // still has wedges has its next four in the current row
                    // (one w, four consecutive u) -- nearly always where the
                    // rows are long, which is where the wedges are: one 16-byte
                    // load of the four ids, w only when its position changed,
                    // no per-wedge stepping.  Waves with short rows stay on the
                    // general path, so they never run both.
                    if (k0 + 4 <= kn && __ballot(left != 0 && (left < 4 || d - u_i < 4)) == 0) {
                        uint2 *const dst = lwu + __umul24(lane, pad) + k0;
                        if (left == 0) {  // a finished lane: empty slots
                            dst[0] = dst[1] = dst[2] = dst[3] = make_uint2(kSentinel, 0u);
                            continue;
                        }
                        const OffT wp = vb + w_i;
                        if (wp != last_wpos) {
                            last_wv = colInd[wp];
                            last_wpos = wp;
                        }
                        const u32x4_e q = *(gptr<u32x4_e>)(colInd + (vb + u_i));
                        dst[0] = make_uint2(last_wv, q.x);
                        dst[1] = make_uint2(last_wv, q.y);
                        dst[2] = make_uint2(last_wv, q.z);
                        dst[3] = make_uint2(last_wv, q.w);
                        left -= 4;
                        u_i += 4;
                        if (left && u_i >= d) {  // the row ended with the group
                            w_i++;
                            u_i = w_i + 1;
                            if (u_i >= d) {  // vertex exhausted: next vertex with wedges
                                const uint64_t i = lend - left;  // the next wedge
                                v = (ws[v + 2] > i) ? v + 1 : rightmost_le(ws, v + 2, v_hi, i);
                                vb = rowPtr[v];
                                d = (uint32_t)(rowPtr[v + 1] - vb);
                                w_i = 0;
                                u_i = 1;
                            }
                        }
                        continue;
                    }
                    // general path: wedge by wedge across row and vertex ends
                    OffT wpos[4], upos[4];
                    bool ok[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        ok[j] = (k0 + j < kn) && left != 0;
                        wpos[j] = 0;
                        upos[j] = 0;
                        if (ok[j]) {
                            wpos[j] = vb + w_i;
                            upos[j] = vb + u_i;
                            left--;
                            if (left) {  // step to the lane's next wedge
                                u_i++;
                                if (u_i >= d) {
                                    w_i++;
                                    u_i = w_i + 1;
                                    if (u_i >= d) {  // vertex exhausted: next vertex with wedges
                                        const uint64_t i = lend - left;  // the next wedge
                                        v = (ws[v + 2] > i) ? v + 1 : rightmost_le(ws, v + 2, v_hi, i);
                                        vb = rowPtr[v];
                                        d = (uint32_t)(rowPtr[v + 1] - vb);
                                        w_i = 0;
                                        u_i = 1;
                                    }
                                }
                            }
                        }
                    }
                    uint32_t wv[4], uv[4], wl[4];
                    // Inside a row a full group's u positions are consecutive:
                    // then one 16-byte load per lane fetches all four (one
                    // vL1D access instead of four).  Across a row end, and for
                    // a lane's last partial group, the ids are loaded one by
                    // one, so no load reaches past the lane's last wedge (nor
                    // past the end of colInd).
                    bool cont = ok[0];
#pragma unroll
                    for (int j = 1; j < 4; j++) cont = cont && ok[j] && upos[j] == upos[0] + j;
#pragma unroll
                    for (int j = 0; j < 4; j++) {  // ok[] is a prefix of the group
                        const OffT prev = j == 0 ? last_wpos : wpos[j - 1];
                        wl[j] = ok[j] && wpos[j] != prev ? colInd[wpos[j]] : 0u;
                    }
                    if (cont) {
                        const u32x4_e q = *(gptr<u32x4_e>)(colInd + upos[0]);
                        uv[0] = q.x;
                        uv[1] = q.y;
                        uv[2] = q.z;
                        uv[3] = q.w;
                    } else {
#pragma unroll
                        for (int j = 0; j < 4; j++) uv[j] = ok[j] ? colInd[upos[j]] : 0u;
                    }
#pragma unroll
                    for (int j = 0; j < 4; j++) uv[j] = ok[j] ? uv[j] : 0u;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        if (ok[j]) {
                            if (wpos[j] != last_wpos) {
                                last_wv = wl[j];
                                last_wpos = wpos[j];
                            }
                            wv[j] = last_wv;
                        } else {
                            wv[j] = kSentinel;
                        }
                    }
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        if (k0 + j < kn) lwu[__umul24(lane, pad) + k0 + j] = make_uint2(wv[j], uv[j]);
                }
                __syncwarp();
                // phase 2: consecutive lanes probe consecutive wedges; each lane
                // advances kWays independent probes in lockstep so that their
                // dependent loads overlap (memory-level parallelism per lane).
                // Step j of lane l reads staged wedge p = 64j + l, which lane
                // p / kn wrote as its (p % kn)-th, at slot (p / kn) * pad + p % kn
                // = p + (p / kn)(pad - kn).  When kn divides 64 that is
                // slot0 + j * (64 / kn) * pad with slot0 the slot of step 0: one
                // add per read.  Otherwise p / kn = (p * magic) >> 16 with
                // magic = ceil(2^16 / kn), exact for p < 64 * kChunkMax = 1024
                // (the error p(magic kn - 2^16) / 2^16 < 1/kn), in 24-bit
                // multiplies.
                const uint32_t magic = (65536u + kn - 1) / kn;
                const uint32_t dpad = pad - kn;
                const bool lin = (kWave % kn) == 0;  // wave-uniform
                const uint32_t lstep = (kWave / kn) * pad;
                const uint2 *const lrow = lwu + (lane + __umul24(__umul24(lane, magic) >> 16, dpad));
                for (uint32_t j0 = 0; j0 < kn; j0 += kWays) {
                    uint32_t pw[kWays], pu[kWays];
#pragma unroll
                    for (int k = 0; k < kWays; k++) {
                        pw[k] = kSentinel;
                        pu[k] = 0;
                        if (j0 + k < kn) {  // wave-uniform
                            uint2 e;
                            if (lin) {
                                e = lrow[(j0 + k) * lstep];
                            } else {
                                const uint32_t p = (j0 + k) * kWave + lane;
                                const uint32_t src = __umul24(p, magic) >> 16;
                                e = lwu[__umul24(src, dpad) + p];
                            }
                            pw[k] = e.x;
                            pu[k] = e.y;
                        }
                    }
                    // Every way's first load is issued before any is used, so
                    // the ways' round trips overlap: a bitmap word for w in the
                    // top-a ranks (tc.cu:262-272), else rowPtr[w], rowPtr[w+1]
                    // for the sorted-list probe of adj+(w) (tc.cu:273-286).
                    // (tc.cu:262-272: w in the top-a ranks, r = w - thr < min(n, a);
                    // empty slots hold w = kSentinel, above every rank)
                    bool isb[kWays], isl[kWays];
                    uint32_t bcol[kWays], bw[kWays];
                    uint64_t bwd[kWays];
                    OffT r0[kWays], r1[kWays];
#pragma unroll
                    for (int k = 0; k < kWays; k++) {
                        const uint32_t r = pw[k] - thr;
                        isb[k] = r < bm_rows;
                        isl[k] = pw[k] < thr;
                        bcol[k] = pu[k] - thr;
                        // word of (w, u) in the square layout: a 24-bit
                        // multiply-add while the bitmap has < 2^32 words
                        bwd[k] = bm_narrow ? (uint64_t)(__umul24(r, a32) + (bcol[k] >> 5))
                                           : (uint64_t)r * a32 + (bcol[k] >> 5);
                    }
#pragma unroll
                    for (int k = 0; k < kWays; k++) bw[k] = isb[k] ? bitmap[bwd[k]] : 0u;
#pragma unroll
                    for (int k = 0; k < kWays; k++) {
                        r0[k] = 0;
                        r1[k] = 0;
                        if (isl[k]) {
                            const gptr<OffT> rp = rowPtr + pw[k];
                            r0[k] = rp[0];
                            r1[k] = rp[1];
                        }
                    }
                    // Sorted-list probe of u in adj+(w) = colInd[r0, r1): a
                    // halving search over [r0 - 1, r1) whose first position
                    // stands for an element below u and is never read.  base
                    // is a position below u's lower bound and base + len one
                    // at or above it (r1: past the row); a step reads
                    // base + len/2, moves base there if that element is below
                    // u, and len loses len/2.  The lower bound, when it lies
                    // in the row, is the last position read at or above u, so
                    // a hit shows as an equal element read at some step; the
                    // hits of a way are kept as a lane mask (one compare per
                    // step, no per-step count).  At len = 1 a probe is done:
                    // len/2 = 0 then moves nothing, so finished, bitmap and
                    // empty slots need no masking beyond their loads, and a
                    // stale element only repeats a hit already in the mask.
                    OffT bs[kWays];
                    uint32_t len[kWays], x[kWays];
                    uint64_t hit[kWays];
#pragma unroll
                    for (int k = 0; k < kWays; k++) {
                        tile_cnt += (uint64_t)__popcll(__ballot(__builtin_amdgcn_ubfe(bw[k], bcol[k], 1u)));
                        bs[k] = r0[k] - 1;
                        len[k] = isl[k] ? (uint32_t)(r1[k] - r0[k]) + 1u : 1u;
                        x[k] = kSentinel;  // never a vertex id
                        hit[k] = 0;
                    }
                    for (;;) {
                        bool more = false;
#pragma unroll
                        for (int k = 0; k < kWays; k++) more |= len[k] > 1u;
                        // wave-uniform exit: every lane stays in the loop until
                        // all are done, so the ballots below see the whole wave
                        if (__ballot(more) == 0) break;
#pragma unroll
                        for (int k = 0; k < kWays; k++)
                            if (len[k] > 1u) x[k] = colInd[bs[k] + (len[k] >> 1)];
#pragma unroll
                        for (int k = 0; k < kWays; k++) {
                            const uint32_t half = len[k] >> 1;
                            hit[k] |= __ballot(x[k] == pu[k]);
                            bs[k] = x[k] < pu[k] ? bs[k] + half : bs[k];
                            len[k] -= half;
                        }
                    }
#pragma unroll
                    for (int k = 0; k < kWays; k++) tile_cnt += (uint64_t)__popcll(hit[k]);
                }
                __syncwarp();
            }

            if (g != cur_g) {
                if (cur_cnt != 0 && lane == 0) atomicAdd(&counts[cur_g], (unsigned long long)cur_cnt);
                cur_g = g;
                cur_cnt = 0;
            }
            cur_cnt += tile_cnt;
        }
    }

    // one 64-bit atomic per block and graph
    if (lane == 0) {
        s_g[wid] = cur_g;
        s_c[wid] = cur_cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t a = 0; a < wpb; a++) {
            if (s_g[a] == kSentinel || s_c[a] == 0) continue;
            uint64_t sum = s_c[a];
            for (uint32_t b = a + 1; b < wpb; b++)
                if (s_g[b] == s_g[a]) {
                    sum += s_c[b];
                    s_c[b] = 0;
                }
            atomicAdd(&counts[s_g[a]], (unsigned long long)sum);
        }
    }
}

// ------------------------------------------------------------- row walk

// The last kTail entries of every row form the row's short triangle; rows
// with kTail or more u after w are its long rows.
constexpr uint32_t kTail = 64;
constexpr uint32_t kTailPairs = kTail * (kTail - 1) / 2;  // 2016

// Wedges of a vertex with out-degree d in its long rows: the first
// C(d,2) - C(kTail,2) wedges of its row-major order.
__device__ __forceinline__ uint64_t long_wedges(uint32_t d) {
    return d > kTail ? binom2_u32(d) - (uint64_t)kTailPairs : 0ull;
}

// k_walk counts the same wedges in the same order as k_count (the order of
// tc.cu:203-240: vertices ascending, each vertex's wedges (w_i, u_i)
// row-major), but assigns them to lanes differently.  A wave owns a
// contiguous range of the order (the tiles of a grab) and walks it in steps
// of up to 64 consecutive wedges, one per lane, so consecutive lanes always
// probe consecutive wedges (the cooperative probe order of tc.cu:251-287)
// without per-lane decoding or LDS staging:
//  * a long step covers up to 64 wedges of the long rows of one vertex:
//    lane l takes u at position ui + l of row wi, or of row wi + 1 when row
//    wi ends inside the step (rows are at least 64 long, so a step spans at
//    most two); the u are one coalesced load and w is one of two
//    wave-uniform ids;
//  * a tail step covers up to 64 wedges of one vertex's short triangle (the
//    pairs of its last m = min(d, 64) entries): the m entries are loaded
//    once into lanes 64 - m .. 63 of a register, and lane l takes pair
//    p0 + l of the 64-entry triangle (the short triangle's pairs are its
//    last C(m,2) pairs in row-major order) from an LDS table of (i, j), then
//    w and u by two shuffles.
// Only the walk state (vertex, row, position) is per wave; it advances on
// the scalar unit.  kSteps steps are prepared before any of their loads is
// used, so a lane has kSteps independent probes in flight.
template <typename OffT, int kSteps>
__global__ __launch_bounds__(kWave * kCountWavesMax) void k_walk(const GraphDesc *__restrict__ desc, uint32_t G,
                                                                 const uint64_t *__restrict__ tree,
                                                                 uint64_t *__restrict__ meta,
                                                                 unsigned long long *__restrict__ counts,
                                                                 uint32_t work) {
    __shared__ uint16_t s_pair[kTailPairs];  // pair p of the 64-triangle: i | j << 8
    __shared__ uint32_t s_g[kCountWavesMax];
    __shared__ uint64_t s_c[kCountWavesMax];
    for (uint32_t i = threadIdx.x; i < kTail - 1; i += blockDim.x) {
        const uint32_t r0 = i * (2 * kTail - 1 - i) / 2;  // first pair of row i
        for (uint32_t j = i + 1; j < kTail; j++) s_pair[r0 + j - i - 1] = (uint16_t)(i | (j << 8));
    }
    __syncthreads();

    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = uni32(threadIdx.x >> 6);
    const uint32_t wpb = blockDim.x >> 6;

    // work distribution as in k_count: XCD-interleaved super-chunks of
    // grabs of consecutive tiles from the XCD's own counter
    const uint64_t ntiles = meta[kMetaTiles];
    const uint64_t waves = (uint64_t)gridDim.x * wpb;
    uint64_t tpg = ntiles / (2 * waves);
    const uint64_t gmax = (work & 0xffffu) ? (work & 0xffffu) : kTilesPerGrabMax;
    tpg = tpg < 1 ? 1 : tpg > gmax ? gmax : tpg;
    const uint64_t ngrabs = (ntiles + tpg - 1) / tpg;
    // dealing to the XCDs (work bit 16) or one stream for the whole grid
    const bool deal = (work >> 16) & 1u;
    const uint32_t streams = deal ? 8u : 1u;
    uint32_t lg = 0;
    while (deal && lg < 6 && (ngrabs >> (lg + 1)) >= 8ull * 64) lg++;
    const uint32_t xcd = deal ? blockIdx.x & 7u : 0u;
    unsigned long long *const ctr = reinterpret_cast<unsigned long long *>(meta + kMetaNext + xcd);

    uint32_t cur_g = kSentinel;
    uint64_t cur_cnt = 0;

    for (;;) {
        unsigned long long grab = 0;
        if (lane == 0) grab = atomicAdd(ctr, 1ull);
        const uint64_t gi = uni64(grab);
        const uint64_t chunk_id = xcd + streams * (gi >> lg);
        const uint64_t tb = ((chunk_id << lg) + (gi & ((1ull << lg) - 1))) * tpg;
        if (tb >= ntiles) break;
        const uint64_t te = tb + tpg < ntiles ? tb + tpg : ntiles;

        for (uint64_t t = tb; t < te;) {
            // the grab's tiles of one graph: wedges [pos, pend) of its order
            const uint32_t g = G > 1 ? graph_of(desc, G, t, &GraphDesc::tile_base) : 0u;
            const GraphDesc &D = desc[g];
            const gptr<OffT> rowPtr = to_global(static_cast<const OffT *>(D.rowPtr));
            const gptr<uint32_t> colInd = to_global(D.colInd);
            const gptr<uint32_t> bitmap = to_global(static_cast<const uint32_t *>(D.bitmap));
            // the same arrays for wave-uniform reads (scalar loads)
            const cptr<OffT> rowPtrU = to_const(static_cast<const OffT *>(D.rowPtr));
            const cptr<uint32_t> colIndU = to_const(D.colInd);
            const cptr<uint64_t> wsU = to_const(static_cast<const uint64_t *>(D.ws));
            const uint32_t gn = D.n, vmax = gn - 1, thr = D.thr, bm_rows = D.bm_rows, a32 = D.a >> 5;
            const uint64_t T = (uint64_t)kWave * D.spread;
            const uint64_t wend = D.wbase + D.wcount;
            const uint64_t g_tile_end = D.tile_base + (D.wcount + T - 1) / T;
            const uint64_t t1 = te < g_tile_end ? te : g_tile_end;
            // (the walk state is wave-uniform: readfirstlane keeps it in
            // SGPRs, so the walk runs on the scalar unit)
            uint64_t pos = uni(D.wbase + (t - D.tile_base) * T);
            const uint64_t pe = D.wbase + (t1 - D.tile_base) * T;
            const uint64_t pend = uni(pe < wend ? pe : wend);
            t = t1;
            if (g != cur_g) {
                if (cur_cnt != 0 && lane == 0) atomicAdd(&counts[cur_g], (unsigned long long)cur_cnt);
                cur_g = g;
                cur_cnt = 0;
            }

            // walk state (wave-uniform): vertex v with vb, d, its cv wedges
            // (ws[v+1] - ws[v]: C(d,2), or in a list-part prefix the first
            // wedges of its row-major order) of which the first lv lie in
            // long rows, position iv in them; in the long rows the current
            // row wi, position ui and w id wcur
            uint32_t v = uni((tree != nullptr) ? tree_search(to_const(tree), wsU, gn, pos)
                                               : rightmost_le(wsU, 0u, vmax, pos));
            OffT vb = uni(rowPtrU[v]);
            uint32_t d = uni((uint32_t)(rowPtrU[v + 1] - vb));
            uint64_t wsv = uni(wsU[v]);
            uint64_t cv = uni(wsU[v + 1] - wsv), lv = long_wedges(d), iv = uni(pos - wsv);
            uint32_t wi = 0, ui = 1, wcur = 0;
            if (iv < lv) {
                decode_wedge(d, iv, wi, ui);
                wi = uni32(wi);
                ui = uni32(ui);
                wcur = uni(colIndU[vb + wi]);
            }
            uint64_t hits = 0;
            while (pos < pend) {
                // step parameters (wave-uniform): n wedges; long: u of lanes
                // below `split` at pa + lane, the rest at pb + lane - split,
                // w = wa / wb; tail: window from pa (lanes from `split` =
                // 64 - m valid), pairs from pb
                uint32_t n[kSteps], split[kSteps], wa[kSteps], wb[kSteps];
                bool tail[kSteps];
                OffT pa[kSteps], pb[kSteps];
#pragma unroll
                for (int k = 0; k < kSteps; k++) {
                    n[k] = 0;
                    split[k] = 0;
                    wa[k] = wb[k] = 0;
                    tail[k] = false;
                    pa[k] = pb[k] = 0;
                    if (pos >= pend) continue;
                    if (iv == cv) {  // vertex done: next vertex with wedges (pos = ws[v + 1] < W)
                        v = uni(wsU[v + 2] > pos ? v + 1 : gallop_le(wsU, v + 2, vmax, pos));
                        vb = uni(rowPtrU[v]);
                        d = uni((uint32_t)(rowPtrU[v + 1] - vb));
                        cv = uni(wsU[v + 1] - wsU[v]);
                        lv = long_wedges(d);
                        iv = 0;
                        wi = 0;
                        ui = 1;
                        if (lv) wcur = uni(colIndU[vb]);
                    }
                    // wedges this step may take: up to the range end and
                    // the vertex's last wedge
                    const uint64_t left0 = pend - pos, rem = cv - iv;
                    const uint64_t left = rem < left0 ? rem : left0;
                    if (iv < lv) {
                        uint64_t m = lv - iv;
                        m = m < left ? m : left;
                        const uint32_t na = m < kWave ? (uint32_t)m : kWave;
                        const uint32_t inrow = d - ui;
                        n[k] = na;
                        split[k] = inrow;
                        pa[k] = vb + ui;
                        pb[k] = vb + wi + 2;
                        wa[k] = wcur;
                        wb[k] = uni(na > inrow ? colIndU[vb + wi + 1] : wcur);
                        pos += na;
                        iv += na;
                        ui += na;
                        if (ui >= d) {  // row wi ended: next row (at least 64 long while iv < lv)
                            const uint32_t over = ui - d;
                            wi++;
                            ui = wi + 1 + over;
                            wcur = uni(na > inrow ? wb[k] : iv < lv ? colIndU[vb + wi] : 0u);
                        }
                        pos = uni(pos);
                        iv = uni(iv);
                        wi = uni(wi);
                        ui = uni(ui);
                    } else {
                        const uint32_t mt = d < kTail ? d : kTail;
                        const uint64_t tp = iv - lv;  // < C(mt, 2) <= 2016
                        const uint32_t ct = mt * (mt - 1) / 2;
                        uint64_t m = ct - tp;
                        m = m < left ? m : left;
                        const uint32_t na = m < kWave ? (uint32_t)m : kWave;
                        tail[k] = true;
                        n[k] = na;
                        split[k] = kTail - mt;
                        pa[k] = vb + d - kTail;  // lanes split..63 hold the last mt entries
                        pb[k] = kTailPairs - ct + (uint32_t)tp;
                        pos = uni(pos + na);
                        iv = uni(iv + na);
                    }
                }
                // loads: u of a long step, the short-triangle window of a tail step
                uint32_t ld[kSteps];
#pragma unroll
                for (int k = 0; k < kSteps; k++) {
                    ld[k] = 0;
                    if (!tail[k]) {
                        if (lane < n[k]) ld[k] = colInd[lane < split[k] ? pa[k] + lane : pb[k] + (lane - split[k])];
                    } else if (lane >= split[k]) {
                        ld[k] = colInd[pa[k] + lane];
                    }
                }
                // (w, u) of each lane's wedge; finished lanes get w = kSentinel
                uint32_t pw[kSteps], pu[kSteps];
#pragma unroll
                for (int k = 0; k < kSteps; k++) {
                    const bool act = lane < n[k];
                    if (!tail[k]) {  // wave-uniform
                        pw[k] = lane < split[k] ? wa[k] : wb[k];
                        pu[k] = ld[k];
                    } else {
                        const uint32_t pr = s_pair[(uint32_t)pb[k] + (act ? lane : 0u)];
                        pw[k] = shfl_off(ld[k], pr & 0xffu);
                        pu[k] = shfl_off(ld[k], pr >> 8);
                    }
                    if (!act) {
                        pw[k] = kSentinel;
                        pu[k] = 0;
                    }
                }
                // probes: a bitmap word for w in the top-a ranks
                // (tc.cu:262-272), else the sorted-list probe of adj+(w)
                // (tc.cu:273-286) by the halving search of k_count.  Each
                // way's loads are issued before any result is used; w dies
                // with them, so only u, the bit's column and the search
                // state stay live.
                uint32_t bw[kSteps], bcol[kSteps];
                OffT r0[kSteps], r1[kSteps];
#pragma unroll
                for (int k = 0; k < kSteps; k++) {
                    const uint32_t r = pw[k] - thr;
                    bcol[k] = pu[k] - thr;
                    bw[k] = 0;
                    if (r < bm_rows) bw[k] = bitmap[(uint64_t)r * a32 + (bcol[k] >> 5)];
                    r0[k] = 0;
                    r1[k] = 0;
                    if (pw[k] < thr) {
                        const gptr<OffT> rp = rowPtr + pw[k];
                        r0[k] = rp[0];
                        r1[k] = rp[1];
                    }
                }
                OffT bs[kSteps];
                uint32_t len[kSteps], x[kSteps];
                uint64_t hit[kSteps];
#pragma unroll
                for (int k = 0; k < kSteps; k++) {
                    hits += (uint64_t)__popcll(__ballot(__builtin_amdgcn_ubfe(bw[k], bcol[k], 1u)));
                    bs[k] = r0[k] - 1;
                    len[k] = (uint32_t)(r1[k] - r0[k]) + 1u;  // 1 (nothing to search) unless a list probe
                    x[k] = kSentinel;
                    hit[k] = 0;
                }
                for (;;) {
                    bool more = false;
#pragma unroll
                    for (int k = 0; k < kSteps; k++) more |= len[k] > 1u;
                    if (__ballot(more) == 0) break;
#pragma unroll
                    for (int k = 0; k < kSteps; k++)
                        if (len[k] > 1u) x[k] = colInd[bs[k] + (len[k] >> 1)];
#pragma unroll
                    for (int k = 0; k < kSteps; k++) {
                        const uint32_t half = len[k] >> 1;
                        hit[k] |= __ballot(x[k] == pu[k]);
                        bs[k] = x[k] < pu[k] ? bs[k] + half : bs[k];
                        len[k] -= half;
                    }
                }
#pragma unroll
                for (int k = 0; k < kSteps; k++) hits += (uint64_t)__popcll(hit[k]);
            }
            cur_cnt += hits;
        }
    }

    if (lane == 0) {
        s_g[wid] = cur_g;
        s_c[wid] = cur_cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t a = 0; a < wpb; a++) {
            if (s_g[a] == kSentinel || s_c[a] == 0) continue;
            uint64_t sum = s_c[a];
            for (uint32_t b = a + 1; b < wpb; b++)
                if (s_g[b] == s_g[a]) {
                    sum += s_c[b];
                    s_c[b] = 0;
                }
            atomicAdd(&counts[s_g[a]], (unsigned long long)sum);
        }
    }
}

// ------------------------------------------------------ small-graph count

// Independent wedge streams each lane of k_count_small advances together.
constexpr int kSmallWays = 4;

// The same count for a graph small enough to keep in LDS (the role of the
// vertex- and edge-parallel kernels of experiments/tc_vertex.cu:113-274 and
// experiments/tc_edge.cu:179-352 for small graphs): one block per graph
// copies its rowPtr, wedge prefix and colInd into LDS once, then every probe
// is an LDS read instead of an L1 lookup, and no tile bookkeeping is paid.
// Block b takes part parts[b] of a graph of mode kModeSmall: a slice of its
// wedge range (graphs with many wedges are split over several blocks, each
// staging its own copy).  The CSR is staged with eight loads in flight per
// lane.  The block's wedge slice is cut into kSmallBlock * kSmallWays
// contiguous streams; lane t advances streams t, t + kSmallBlock, ... in
// lockstep so that their LDS round trips overlap.  A stream's first wedge is
// located by a search of the prefix and decoded exactly, the rest by
// stepping.  The probe of u in adj+(w) is a lower-bound search that starts
// from the previous probe's position while w stays (u only grows along a
// row); a present u is always probed once (the last probe at or above u is
// the lower bound), so a hit is counted when seen.  Every loop exits by a
// wave-wide vote, so lanes never wait in nested per-lane loops.  Counts of
// a mode-small graph need 2(n+1) + m <= kSmallWordsMax words of LDS, which
// the host checks against the launch's dynamic LDS size.
template <typename OffT>
__global__ __launch_bounds__(kSmallBlock) void k_count_small(const GraphDesc *__restrict__ desc,
                                                             const SmallPart *__restrict__ parts,
                                                             unsigned long long *__restrict__ counts) {
    extern __shared__ uint32_t lds[];
    const SmallPart sp = parts[blockIdx.x];
    const uint32_t g = sp.g;
    const GraphDesc &D = desc[g];
    // this block's slice of the graph's range [wbase, wbase + wcount); wb + wc <= W < 2^31
    const uint32_t wtot = (uint32_t)D.wcount;
    const uint32_t s0 = (uint32_t)((uint64_t)wtot * sp.part / sp.parts);
    const uint32_t s1 = (uint32_t)((uint64_t)wtot * (sp.part + 1u) / sp.parts);
    if (s1 == s0) return;  // block-uniform
    const uint32_t n = D.n;
    const gptr<OffT> rowPtr = to_global(static_cast<const OffT *>(D.rowPtr));
    const gptr<uint64_t> ws = to_global(static_cast<const uint64_t *>(D.ws));
    const OffT base = rowPtr[0];
    const uint32_t m = (uint32_t)(rowPtr[n] - base);
    const gptr<uint32_t> colInd = to_global(D.colInd) + base;
    uint32_t *const rp = lds;              // n+1 row offsets relative to rowPtr[0]
    uint32_t *const wp = lds + (n + 1);    // n+1 wedge prefix; W <= C(m,2) < 2^31
    uint32_t *const col = lds + 2 * (n + 1);
    constexpr uint32_t kStage = 8;  // loads in flight per lane
    for (uint32_t i0 = 0; i0 <= n; i0 += kStage * kSmallBlock) {
        OffT r[kStage];
        uint64_t x[kStage];
#pragma unroll
        for (uint32_t q = 0; q < kStage; q++) {
            const uint32_t i = i0 + q * kSmallBlock + threadIdx.x;
            r[q] = i <= n ? rowPtr[i] : base;
            x[q] = i <= n ? ws[i] : 0;
        }
#pragma unroll
        for (uint32_t q = 0; q < kStage; q++) {
            const uint32_t i = i0 + q * kSmallBlock + threadIdx.x;
            if (i <= n) {
                rp[i] = (uint32_t)(r[q] - base);
                wp[i] = (uint32_t)x[q];
            }
        }
    }
    for (uint32_t i0 = 0; i0 < m; i0 += kStage * kSmallBlock) {
        uint32_t c[kStage];
#pragma unroll
        for (uint32_t q = 0; q < kStage; q++) {
            const uint32_t i = i0 + q * kSmallBlock + threadIdx.x;
            c[q] = i < m ? colInd[i] : 0u;
        }
#pragma unroll
        for (uint32_t q = 0; q < kStage; q++) {
            const uint32_t i = i0 + q * kSmallBlock + threadIdx.x;
            if (i < m) col[i] = c[q];
        }
    }
    __syncthreads();

    // streams of `per` wedges: stream j = k * kSmallBlock + lane covers
    // [wb + j*per, min(wb + (j+1)*per, wb + wc))
    const uint32_t wb = (uint32_t)D.wbase + s0, wc = s1 - s0;
    const uint32_t per = (wc + kSmallBlock * kSmallWays - 1) / (kSmallBlock * kSmallWays);
    uint32_t i[kSmallWays], iend[kSmallWays], v[kSmallWays], vb[kSmallWays], d[kSmallWays];
    uint32_t w_i[kSmallWays], u_i[kSmallWays], lo[kSmallWays], hi[kSmallWays];
#pragma unroll
    for (int k = 0; k < kSmallWays; k++) {
        const uint32_t off = (k * kSmallBlock + threadIdx.x) * per;  // < kSmallBlock*kSmallWays*per < 2^32
        i[k] = iend[k] = 0;
        v[k] = vb[k] = d[k] = w_i[k] = u_i[k] = lo[k] = hi[k] = 0;
        if (off < wc) {
            i[k] = wb + off;
            iend[k] = wb + (wc - off < per ? wc : off + per);
            v[k] = rightmost_le(wp, 0u, n - 1, i[k]);
            vb[k] = rp[v[k]];
            d[k] = rp[v[k] + 1] - vb[k];
            decode_wedge(d[k], i[k] - wp[v[k]], w_i[k], u_i[k]);
            const uint32_t w = col[vb[k] + w_i[k]];
            lo[k] = rp[w];
            hi[k] = rp[w + 1];
        }
    }
    uint32_t cnt = 0;
    for (;;) {
        bool act[kSmallWays], any = false;
#pragma unroll
        for (int k = 0; k < kSmallWays; k++) {
            act[k] = i[k] < iend[k];
            any |= act[k];
        }
        if (__ballot(any) == 0) break;
        uint32_t u[kSmallWays], len[kSmallWays];
#pragma unroll
        for (int k = 0; k < kSmallWays; k++) {
            u[k] = act[k] ? col[vb[k] + u_i[k]] : 0u;
            len[k] = act[k] ? hi[k] - lo[k] : 0u;
        }
        for (;;) {
            bool more = false;
#pragma unroll
            for (int k = 0; k < kSmallWays; k++) more |= len[k] > 0;
            if (__ballot(more) == 0) break;
            uint32_t x[kSmallWays];
#pragma unroll
            for (int k = 0; k < kSmallWays; k++) x[k] = len[k] > 0 ? col[lo[k] + (len[k] >> 1)] : 0u;
#pragma unroll
            for (int k = 0; k < kSmallWays; k++) {
                const uint32_t half = len[k] >> 1;
                const bool lt = x[k] < u[k];
                const uint32_t go = len[k] > 0 && lt ? half + 1 : 0u;
                cnt += (len[k] > 0 && x[k] == u[k]) ? 1u : 0u;
                lo[k] += go;
                len[k] = lt ? len[k] - go : half;
            }
        }
        // step every active stream to its next wedge
#pragma unroll
        for (int k = 0; k < kSmallWays; k++) {
            const bool next = act[k] && ++i[k] < iend[k];
            const bool new_w = next && ++u_i[k] == d[k];
            if (new_w) {
                w_i[k]++;
                u_i[k] = w_i[k] + 1;
                if (u_i[k] == d[k]) {
                    // a wedge i < W remains, so a later vertex holds it: v + 2 <= n
                    v[k] = (wp[v[k] + 2] > i[k]) ? v[k] + 1 : rightmost_le(wp, v[k] + 2, n - 1, i[k]);
                    vb[k] = rp[v[k]];
                    d[k] = rp[v[k] + 1] - vb[k];
                    w_i[k] = 0;
                    u_i[k] = 1;
                }
                const uint32_t w = col[vb[k] + w_i[k]];
                lo[k] = rp[w];
                hi[k] = rp[w + 1];
            }
        }
    }
    const uint64_t tot = wave_sum((uint64_t)cnt);
    if ((threadIdx.x & (kWave - 1)) == 0 && tot != 0) atomicAdd(&counts[g], (unsigned long long)tot);
}

// Plain streaming copy (16-byte vectors), the reference kernel for scaling
// the HBM byte counters of a profiler run.
__global__ __launch_bounds__(kBlock) void k_copy_u32x4(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                       uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// ---------------------------------------------------------------- launches

// Launches KERNEL<uint64_t> or KERNEL<uint32_t> by row-offset width.


hipError_t launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t st) {
    const uint64_t n = bytes / 16;
    if (n == 0) return hipSuccess;
    k_copy_u32x4<<<dim3(4096), dim3(kBlock), 0, st>>>(static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n);
    return hipGetLastError();
}

hipError_t launch_fill(uint32_t *p, uint64_t words, uint32_t v, hipStream_t st) {
    if (words == 0) return hipSuccess;
    const uint64_t fb = (words + kBlock - 1) / kBlock;
    k_fill_u32<<<dim3((uint32_t)(fb < 4096 ? fb : 4096)), dim3(kBlock), 0, st>>>(p, words, v);
    return hipGetLastError();
}

hipError_t launch_debug_decode(const uint32_t *d, const uint64_t *iv, uint64_t k, uint32_t *w, uint32_t *u,
                               hipStream_t st) {
    if (k == 0) return hipSuccess;
    k_debug_decode<<<dim3((uint32_t)((k + kBlock - 1) / kBlock)), dim3(kBlock), 0, st>>>(d, iv, k, w, u);
    return hipGetLastError();
}

hipError_t launch_scan_partials(uint64_t *partials, uint32_t nb, hipStream_t st) {
    if (nb == 0) return hipSuccess;
    k_scan_partials<<<dim3(1), dim3(kBlock), 0, st>>>(partials, nb);
    return hipGetLastError();
}

uint32_t prefix_blocks(uint32_t n) { return (uint32_t)(((uint64_t)n + kScanChunk - 1) / kScanChunk); }

hipError_t launch_prefix(const GraphDesc *desc, uint32_t G, int ob, uint32_t nblocks, uint64_t *partials,
                         hipStream_t st) {
    if (nblocks == 0) return hipSuccess;
    hipError_t e = WTC_LAUNCH_OFF(ob, k_wedge_partials, dim3(nblocks), dim3(kBlock), 0, st, desc, G, partials);
    if (e != hipSuccess) return e;
    k_scan_partials<<<dim3(1), dim3(kBlock), 0, st>>>(partials, nblocks);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return WTC_LAUNCH_OFF(ob, k_wedge_scan, dim3(nblocks), dim3(kBlock), 0, st, desc, G,
                          (const uint64_t *)partials);
}

hipError_t launch_setup(GraphDesc *desc, uint32_t G, uint64_t *tree, uint64_t *meta, unsigned long long *counts,
                        hipStream_t st) {
    k_setup<<<dim3(1), dim3(kBlock), 0, st>>>(desc, G, tree, meta, counts);
    return hipGetLastError();
}

hipError_t launch_bitmap(const GraphDesc *desc, uint32_t G, int ob, uint64_t rows, uint64_t clear_words,
                         hipStream_t st) {
    if (rows == 0) return hipSuccess;
    if (clear_words) {
        const uint64_t cb = (clear_words + kBlock - 1) / kBlock;
        k_bitmap_clear<<<dim3((uint32_t)(cb < 4096 ? cb : 4096)), dim3(kBlock), 0, st>>>(desc, G, clear_words);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const uint64_t bb = (rows + kWavesPerBlock - 1) / kWavesPerBlock;
    return WTC_LAUNCH_OFF(ob, k_bitmap_build, dim3((uint32_t)(bb < 8192 ? bb : 8192)), dim3(kBlock), 0, st, desc, G,
                          rows);
}

hipError_t launch_count_small(const GraphDesc *desc, const SmallPart *parts, uint32_t nparts, int ob,
                              unsigned long long *counts, uint32_t lds_bytes, hipStream_t st) {
    if (nparts == 0) return hipSuccess;
    if (lds_bytes == 0 || lds_bytes > kSmallWordsMax * 4u) return hipErrorInvalidValue;
    return WTC_LAUNCH_OFF(ob, k_count_small, dim3(nparts), dim3(kSmallBlock), lds_bytes, st, desc, parts, counts);
}

hipError_t allow_small_lds(uint32_t lds_bytes) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_count_small<uint32_t>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_count_small<uint64_t>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
}

size_t count_lds_bytes(uint32_t max_spread, int wpb) {
    const uint32_t sc = max_spread < kChunkMax ? max_spread : kChunkMax;
    const uint32_t pad = sc | 1u;
    return (size_t)wpb * kWave * pad * sizeof(uint2);
}

static uint32_t count_pad(uint32_t max_spread) {
    const uint32_t sc = max_spread < kChunkMax ? max_spread : kChunkMax;
    return sc | 1u;
}

// Calls f.template run<W>() for probe_ways W.
template <class F>
static hipError_t with_ways(int ways, F &f) {
    switch (ways) {
        case 1: return f.template run<1>();
        case 2: return f.template run<2>();
        case 8: return f.template run<8>();
        default: return f.template run<4>();
    }
}

struct OccupancyQuery {
    int ob, wpb, kernel;
    size_t shm;
    int *per_cu;
    template <int W>
    hipError_t run() {
        const int bs = wpb * kWave;
        if (kernel == kKernelWalk)
            return ob == 8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_walk<uint64_t, W>, bs, 0)
                           : hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_walk<uint32_t, W>, bs, 0);
        return ob == 8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_count<uint64_t, W>, bs, shm)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_count<uint32_t, W>, bs, shm);
    }
};

bool valid_count_waves(int wpb) { return wpb == 1 || wpb == 2 || wpb == 4 || wpb == 8; }

hipError_t count_grid(int kernel, int ob, uint32_t max_spread, int ways, int wpb, int num_cus, uint32_t *grid) {
    if (!valid_count_waves(wpb)) return hipErrorInvalidValue;
    int per_cu = 0;
    OccupancyQuery q{ob, wpb, kernel, count_lds_bytes(max_spread, wpb), &per_cu};
    hipError_t e = with_ways(ways, q);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    uint64_t g = (uint64_t)per_cu * (uint64_t)(num_cus > 0 ? num_cus : 1);
    g = (g / 8) * 8;
    if (g < 8) g = 8;
    *grid = (uint32_t)g;
    return hipSuccess;
}

struct CountLaunch {
    const GraphDesc *desc;
    uint32_t G;
    int ob, wpb, kernel;
    uint32_t work;
    const uint64_t *tree;
    uint64_t *meta;
    unsigned long long *counts;
    uint32_t pad, chunk;
    size_t shm;
    uint32_t grid;
    hipStream_t st;
    template <int W>
    hipError_t run() {
        const dim3 blk(wpb * kWave);
        if (kernel == kKernelWalk) {
            if (ob == 8) k_walk<uint64_t, W><<<dim3(grid), blk, 0, st>>>(desc, G, tree, meta, counts, work);
            else k_walk<uint32_t, W><<<dim3(grid), blk, 0, st>>>(desc, G, tree, meta, counts, work);
        } else if (ob == 8) {
            k_count<uint64_t, W><<<dim3(grid), blk, shm, st>>>(desc, G, tree, meta, counts, pad, chunk, work);
        } else {
            k_count<uint32_t, W><<<dim3(grid), blk, shm, st>>>(desc, G, tree, meta, counts, pad, chunk, work);
        }
        return hipGetLastError();
    }
};

hipError_t launch_count(int kernel, const GraphDesc *desc, uint32_t G, int ob, const uint64_t *tree, uint64_t *meta,
                        unsigned long long *counts, uint32_t max_spread, uint32_t chunk, int ways, int wpb,
                        uint32_t work, uint32_t grid, hipStream_t st) {
    if (chunk < 1 || chunk > kChunkMax || !valid_count_waves(wpb)) return hipErrorInvalidValue;
    if (kernel != kKernelWalk && kernel != kKernelStaged) return hipErrorInvalidValue;
    const uint32_t eff = max_spread < chunk ? max_spread : chunk;
    CountLaunch l{desc, G, ob, wpb, kernel, work, tree, meta, counts, count_pad(eff), chunk,
                  count_lds_bytes(eff, wpb), grid, st};
    return with_ways(ways, l);
}

}  // namespace wtc
