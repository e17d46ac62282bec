// wtc_kernels.hip -- HIP kernels of the wedge-parallel triangle counter for
// CDNA4 (gfx950, wave64).
//
// Device pipeline for a graph set (one graph or a batch laid out as a
// disjoint union, see GraphDesc):
//   k_wedge_partials / k_scan_partials / k_wedge_scan
//       wedgeSum[v] = sum_{x<v} C(d+(x),2) over the union's vertices
//       (the host prefix of tc.cu:348-361, as a reduce-then-scan in HBM)
//   k_setup       per-graph wedge ranges, tile bases, the 12-level search
//                 tree over wedgeSum (tc.cu:306-313) and zeroed counts
//   k_fill_u32 / k_bitmap_build
//                 the strict-upper-triangular bitmap over each graph's top-a
//                 ranks (tc.cu:329-346), built with atomicOr
//   k_count       the wedge-parallel count (tc.cu:173-293): persistent,
//                 XCD-partitioned grid; every wave owns tiles of 64*s wedges
//                 and keeps its own LDS staging, so waves never wait on each
//                 other inside a tile.
#include "wtc_kernels.h"
#include "wtc_launch.h"

namespace wtc {

__device__ __forceinline__ uint64_t binom2_d(uint64_t x) {
    return (x & 1u) ? x * ((x - 1u) >> 1) : (x >> 1) * (x - 1u);
}

__device__ __forceinline__ uint32_t uni32(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t x) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        uint64_t y = __shfl_up(x, (unsigned)o, kWave);
        if (lane >= o) x += y;
    }
    return x;
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t x) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
    return x;
}

// Exclusive scan over the 256 lanes of a block; s_tot holds 4 wave totals.
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t x, uint64_t *s_tot, uint64_t &total) {
    const uint64_t inc = wave_incl_scan(x);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & (kWave - 1)) == kWave - 1) s_tot[w] = inc;
    __syncthreads();
    uint64_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; i++) {
        const uint64_t t = s_tot[i];
        off += (i < w) ? t : 0;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return off + inc - x;
}

// ------------------------------------------------------------ wedge prefix

template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_wedge_partials(const OffT *__restrict__ rowPtr, uint32_t N,
                                                           uint64_t *__restrict__ partials) {
    __shared__ uint64_t s_tot[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    uint64_t acc = 0;
#pragma unroll 4
    for (int k = 0; k < kScanItems; k++) {
        const uint64_t v = base + (uint64_t)k * kBlock + threadIdx.x;
        if (v < N) acc += binom2_d((uint64_t)(rowPtr[v + 1] - rowPtr[v]));
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
__global__ __launch_bounds__(kBlock) void k_wedge_scan(const OffT *__restrict__ rowPtr, uint32_t N,
                                                       const uint64_t *__restrict__ partials,
                                                       uint64_t *__restrict__ ws) {
    __shared__ uint64_t s_tot[kWavesPerBlock];
    uint64_t carry = partials[blockIdx.x];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    for (int k = 0; k < kScanItems; k++) {
        const uint64_t v = base + (uint64_t)k * kBlock + threadIdx.x;
        const uint64_t x = v < N ? binom2_d((uint64_t)(rowPtr[v + 1] - rowPtr[v])) : 0;
        uint64_t tot;
        const uint64_t ex = block_excl_scan(x, s_tot, tot);
        if (v < N) ws[v] = carry + ex;
        carry += tot;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) ws[N] = carry;
}

// ------------------------------------------------------------------- setup

// One block.  Per graph: wedge range, tile count (64*s wedges per tile) and
// exclusive tile base; meta totals; counts zeroed; search tree for a single
// graph (node i = wedgeSum at the midpoint of its range, tc.cu:306-313).
__global__ __launch_bounds__(kBlock) void k_setup(const uint64_t *__restrict__ ws, GraphDesc *__restrict__ desc,
                                                  uint32_t G, uint64_t *__restrict__ tree,
                                                  uint64_t *__restrict__ meta,
                                                  unsigned long long *__restrict__ counts) {
    __shared__ uint64_t s_tot[kWavesPerBlock];
    uint64_t carry = 0, wsum = 0;
    for (uint32_t base = 0; base < G; base += kBlock) {
        const uint32_t g = base + threadIdx.x;
        uint64_t tiles = 0, wc = 0;
        if (g < G) {
            GraphDesc &D = desc[g];
            const uint64_t w0 = ws[D.vbase], W = ws[(uint64_t)D.vbase + D.n] - w0;
            const uint64_t b = D.w_begin < W ? D.w_begin : W;
            const uint64_t e = D.w_end < W ? D.w_end : W;
            wc = e > b ? e - b : 0;
            D.wbase = w0 + b;
            D.wcount = wc;
            const uint64_t per = (uint64_t)kWave * D.spread;
            tiles = (wc + per - 1) / per;
            counts[g] = 0;
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
    if (tree != nullptr) {  // only for G == 1, where vbase == 0
        const uint32_t n = desc[0].n;
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

template <typename T>
__device__ __forceinline__ uint32_t rightmost_le_desc(const GraphDesc *__restrict__ desc, uint32_t G, T key,
                                                      uint64_t GraphDesc::*field) {
    uint32_t lo = 0, hi = G - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (desc[mid].*field <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One wave per bitmap row (rows v in [n - min(n,a), n) of every graph).
template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_bitmap_build(const OffT *__restrict__ rowPtr,
                                                         const uint32_t *__restrict__ colInd,
                                                         const GraphDesc *__restrict__ desc, uint32_t G,
                                                         uint64_t rows, uint32_t *__restrict__ bm) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t gw = (uint64_t)blockIdx.x * kWavesPerBlock + uni32(threadIdx.x >> 6);
    const uint64_t nw = (uint64_t)gridDim.x * kWavesPerBlock;
    for (uint64_t r = gw; r < rows; r += nw) {
        const uint32_t g = G > 1 ? rightmost_le_desc(desc, G, r, &GraphDesc::row_base) : 0u;
        const GraphDesc &D = desc[g];
        const uint32_t vl = D.n - D.bm_rows + (uint32_t)(r - D.row_base);
        const uint64_t v = (uint64_t)D.vbase + vl;
        const uint64_t rowbits = D.bm_size - binom2_d((uint64_t)D.n - vl) - vl - 1;
        const OffT b = rowPtr[v], e = rowPtr[v + 1];
        for (OffT j = b + lane; j < e; j += kWave) {
            const uint64_t idx = rowbits + (colInd[j] - D.vbase);
            atomicOr(&bm[D.bm_word_off + (idx >> 5)], 1u << (idx & 31u));
        }
    }
}

// ------------------------------------------------------------------- count

// Rightmost v in [lo, hi] with ws[v] <= t, given ws[lo] <= t.  Called with
// wave-uniform arguments (scalar loads) and per lane (vector loads).
__device__ __forceinline__ uint32_t rightmost_le(const uint64_t *__restrict__ ws, uint32_t lo, uint32_t hi,
                                                 uint64_t t) {
    uint32_t s = lo + 1, e = hi + 1;
    while (s < e) {
        const uint32_t mid = s + ((e - s) >> 1);
        if (ws[mid] <= t) s = mid + 1;
        else e = mid;
    }
    return s - 1;
}

// Wave-uniform search for the tile's first vertex over [0, n) of a single
// graph: 12 levels from the compact search tree, then wedgeSum
// (tc.cu:139-170 semantics).
__device__ __forceinline__ uint32_t tree_search(const uint64_t *__restrict__ tree,
                                                const uint64_t *__restrict__ ws, uint32_t n, uint64_t t) {
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
__device__ __forceinline__ uint32_t gallop_le(const uint64_t *__restrict__ ws, uint32_t lo, uint32_t vmax,
                                              uint64_t t) {
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

template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_count(const OffT *__restrict__ rowPtr,
                                                  const uint32_t *__restrict__ colInd,
                                                  const uint64_t *__restrict__ ws,
                                                  const uint64_t *__restrict__ tree,
                                                  const GraphDesc *__restrict__ desc, uint32_t G,
                                                  const uint64_t *__restrict__ meta,
                                                  const uint32_t *__restrict__ bitmap,
                                                  unsigned long long *__restrict__ counts, uint32_t pad_max) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_g[kWavesPerBlock];
    __shared__ uint64_t s_c[kWavesPerBlock];

    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = uni32(threadIdx.x >> 6);
    uint32_t *const lw = lds + wid * (2u * kWave * pad_max);
    uint32_t *const lu = lw + kWave * pad_max;

    // XCD partition: blocks go round-robin to the 8 XCDs, so blocks with the
    // same blockIdx % 8 share an L2; each such group sweeps one contiguous
    // eighth of the tiles.
    const uint64_t ntiles = meta[kMetaTiles];
    const uint32_t xcd = blockIdx.x & 7u, jx = blockIdx.x >> 3, bpx = gridDim.x >> 3;
    const uint64_t t_lo = ntiles * xcd / 8, t_hi = ntiles * (xcd + 1) / 8;
    const uint64_t wpx = (uint64_t)bpx * kWavesPerBlock;

    uint32_t cur_g = kSentinel;
    uint64_t cur_cnt = 0;

    for (uint64_t t = t_lo + (uint64_t)jx * kWavesPerBlock + wid; t < t_hi; t += wpx) {
        const uint32_t g = G > 1 ? rightmost_le_desc(desc, G, t, &GraphDesc::tile_base) : 0u;
        const GraphDesc &D = desc[g];
        const uint32_t s = D.spread;
        const uint32_t sc = s < kChunkMax ? s : kChunkMax;
        const uint32_t pad = sc | 1u;
        const uint64_t wend = D.wbase + D.wcount;
        const uint64_t tile_start = D.wbase + (t - D.tile_base) * ((uint64_t)kWave * s);
        const uint64_t tile_end = tile_start + (uint64_t)kWave * s < wend ? tile_start + (uint64_t)kWave * s : wend;
        const uint32_t vmax = D.vbase + D.n - 1;
        const uint32_t thr = D.thr, vbase = D.vbase, gn = D.n;
        const uint64_t bm_size = D.bm_size, bm_off = D.bm_word_off;

        // wave-uniform vertex range [v_lo, v_hi] of the tile
        const uint32_t v_lo = (tree != nullptr) ? tree_search(tree, ws, gn, tile_start)
                                                : rightmost_le(ws, D.vbase, vmax, tile_start);
        const uint32_t v_hi = gallop_le(ws, v_lo, vmax, tile_end - 1);

        // lane state: wedges [first, lend)
        const uint64_t first = tile_start + (uint64_t)lane * s;
        const uint64_t lend = first + s < tile_end ? first + s : tile_end;
        uint64_t i = first;
        bool started = false;
        uint32_t v = 0, d = 0, w_i = 0, u_i = 0, w = 0;
        OffT vb = 0;
        uint64_t tile_cnt = 0;

        for (uint64_t c = 0; c < s; c += sc) {  // 64-bit: s may be close to 2^32
            const uint32_t kn = (s - c) < sc ? (uint32_t)(s - c) : sc;
            // phase 1: generate this lane's next kn wedges into LDS
            for (uint32_t k = 0; k < kn; k++, i++) {
                uint32_t w_out = kSentinel, u_out = 0;
                if (i < lend) {
                    if (!started) {
                        started = true;
                        v = rightmost_le(ws, v_lo, v_hi, i);
                        vb = rowPtr[v];
                        d = (uint32_t)(rowPtr[v + 1] - vb);
                        decode_wedge(d, i - ws[v], w_i, u_i);
                        w = colInd[vb + w_i];
                    } else {
                        u_i++;
                        if (u_i >= d) {
                            w_i++;
                            u_i = w_i + 1;
                            if (u_i >= d) {  // vertex exhausted: next vertex with wedges
                                v = (ws[v + 2] > i) ? v + 1 : rightmost_le(ws, v + 2, v_hi, i);
                                vb = rowPtr[v];
                                d = (uint32_t)(rowPtr[v + 1] - vb);
                                w_i = 0;
                                u_i = 1;
                            }
                            w = colInd[vb + w_i];
                        }
                    }
                    w_out = w;
                    u_out = colInd[vb + u_i];
                }
                lw[lane * pad + k] = w_out;
                lu[lane * pad + k] = u_out;
            }
            __syncwarp();
            // phase 2: consecutive lanes probe consecutive wedges
            const uint32_t magic = (65536u + kn - 1) / kn;
            for (uint32_t j = 0; j < kn; j++) {
                const uint32_t p = j * kWave + lane;
                const uint32_t q = (p * magic) >> 16;
                const uint32_t r = p - q * kn;
                const uint32_t pw = lw[q * pad + r], pu = lu[q * pad + r];
                bool hit = false;
                if (pw != kSentinel) {
                    if (pw >= thr) {
                        const uint32_t wl = pw - vbase, ul = pu - vbase;
                        const uint64_t idx = bm_size - binom2_d((uint64_t)gn - wl) + (ul - wl - 1);
                        hit = (bitmap[bm_off + (idx >> 5)] >> (idx & 31u)) & 1u;
                    } else {
                        OffT b = rowPtr[pw];
                        OffT len = rowPtr[pw + 1] - b;
                        while (len > 0) {
                            const OffT half = len >> 1;
                            const uint32_t x = colInd[b + half];
                            if (x == pu) {
                                hit = true;
                                break;
                            }
                            if (x < pu) {
                                b += half + 1;
                                len -= half + 1;
                            } else {
                                len = half;
                            }
                        }
                    }
                }
                tile_cnt += (uint64_t)__popcll(__ballot(hit));
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

    // one 64-bit atomic per block and graph
    if (lane == 0) {
        s_g[wid] = cur_g;
        s_c[wid] = cur_cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int a = 0; a < kWavesPerBlock; a++) {
            if (s_g[a] == kSentinel || s_c[a] == 0) continue;
            uint64_t sum = s_c[a];
            for (int b = a + 1; b < kWavesPerBlock; b++)
                if (s_g[b] == s_g[a]) {
                    sum += s_c[b];
                    s_c[b] = 0;
                }
            atomicAdd(&counts[s_g[a]], (unsigned long long)sum);
        }
    }
}

// ---------------------------------------------------------------- launches

// Launches KERNEL<uint64_t> or KERNEL<uint32_t> by row-offset width.
#define WTC_LAUNCH_OFF(ob, KERNEL, GRID, BLOCK, SHMEM, STREAM, ROWPTR, ...)                                      \
    [&]() -> hipError_t {                                                                                        \
        if ((ob) == 8)                                                                                           \
            KERNEL<uint64_t><<<GRID, BLOCK, SHMEM, STREAM>>>(static_cast<const uint64_t *>(ROWPTR), __VA_ARGS__); \
        else                                                                                                     \
            KERNEL<uint32_t><<<GRID, BLOCK, SHMEM, STREAM>>>(static_cast<const uint32_t *>(ROWPTR), __VA_ARGS__); \
        return hipGetLastError();                                                                                \
    }()

hipError_t launch_debug_decode(const uint32_t *d, const uint64_t *iv, uint64_t k, uint32_t *w, uint32_t *u,
                               hipStream_t st) {
    if (k == 0) return hipSuccess;
    k_debug_decode<<<dim3((uint32_t)((k + kBlock - 1) / kBlock)), dim3(kBlock), 0, st>>>(d, iv, k, w, u);
    return hipGetLastError();
}

uint32_t prefix_blocks(uint32_t N) { return (uint32_t)(((uint64_t)N + kScanChunk - 1) / kScanChunk); }

hipError_t launch_prefix(const void *rowPtr, int ob, uint32_t N, uint64_t *partials, uint64_t *ws,
                         hipStream_t st) {
    const uint32_t nb = prefix_blocks(N);
    if (nb == 0) return hipSuccess;  // callers never pass an empty vertex set
    hipError_t e = WTC_LAUNCH_OFF(ob, k_wedge_partials, dim3(nb), dim3(kBlock), 0, st, rowPtr, N, partials);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kBlock), 0, st, partials, nb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return WTC_LAUNCH_OFF(ob, k_wedge_scan, dim3(nb), dim3(kBlock), 0, st, rowPtr, N,
                          (const uint64_t *)partials, ws);
}

hipError_t launch_setup(const uint64_t *ws, GraphDesc *desc, uint32_t G, uint64_t *tree, uint64_t *meta,
                        unsigned long long *counts, hipStream_t st) {
    hipLaunchKernelGGL(k_setup, dim3(1), dim3(kBlock), 0, st, ws, desc, G, tree, meta, counts);
    return hipGetLastError();
}

hipError_t launch_bitmap(const void *rowPtr, int ob, const uint32_t *colInd, const GraphDesc *desc, uint32_t G,
                         uint64_t rows, uint32_t *bm, uint64_t words, hipStream_t st) {
    if (words == 0 || rows == 0) return hipSuccess;
    uint64_t fb = (words + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)(fb < 4096 ? fb : 4096)), dim3(kBlock), 0, st, bm, words, 0u);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    uint64_t bb = (rows + kWavesPerBlock - 1) / kWavesPerBlock;
    return WTC_LAUNCH_OFF(ob, k_bitmap_build, dim3((uint32_t)(bb < 8192 ? bb : 8192)), dim3(kBlock), 0, st, rowPtr,
                          colInd, desc, G, rows, bm);
}

size_t count_lds_bytes(uint32_t max_spread) {
    const uint32_t sc = max_spread < kChunkMax ? max_spread : kChunkMax;
    const uint32_t pad = sc | 1u;
    return (size_t)kWavesPerBlock * 2u * kWave * pad * sizeof(uint32_t);
}

uint32_t count_pad(uint32_t max_spread) {
    const uint32_t sc = max_spread < kChunkMax ? max_spread : kChunkMax;
    return sc | 1u;
}

hipError_t count_grid(int ob, uint32_t max_spread, int num_cus, uint32_t *grid) {
    int per_cu = 0;
    const size_t shm = count_lds_bytes(max_spread);
    hipError_t e = ob == 8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_count<uint64_t>, kBlock, shm)
                           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_count<uint32_t>, kBlock, shm);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    uint64_t g = (uint64_t)per_cu * (uint64_t)(num_cus > 0 ? num_cus : 1);
    g = (g / 8) * 8;
    if (g < 8) g = 8;
    *grid = (uint32_t)g;
    return hipSuccess;
}

hipError_t launch_count(const void *rowPtr, int ob, const uint32_t *colInd, const uint64_t *ws, const uint64_t *tree,
                        const GraphDesc *desc, uint32_t G, const uint64_t *meta, const uint32_t *bm,
                        unsigned long long *counts, uint32_t max_spread, uint32_t grid, hipStream_t st) {
    const size_t shm = count_lds_bytes(max_spread);
    const uint32_t pad = count_pad(max_spread);
    return WTC_LAUNCH_OFF(ob, k_count, dim3(grid), dim3(kBlock), shm, st, rowPtr, colInd, ws, tree, desc, G, meta, bm,
                          counts, pad);
}

}  // namespace wtc
