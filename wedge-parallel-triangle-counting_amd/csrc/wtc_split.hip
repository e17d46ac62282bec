// wtc_split.hip -- the split count of one graph (counting kernel choice
// kKernelSplit).  Every wedge (v, w, u) of a degree-oriented CSR is counted
// by one of two paths, chosen by its middle vertex w:
//
//  * w among the top-a ranks (w >= thr = n - a): the wedges whose probe the
//    reference answers from its top-a bitmap (tc.cu:262-272).  They are
//    counted grouped by w.  The in-edges v -> w (j the position of w in
//    adj+(v), e the row end, tail colInd[j+1 .. e)) are grouped by w with a
//    counting sort of the top-a columns, built per count.  A block of the
//    row kernel holds adj+(w) as a bitmap of a bits in LDS, and its waves
//    stream the tails of w's in-edges through it, four ids per lane per
//    16-byte load.  Each probe is one LDS read, and the top-a bitmap never
//    exists in HBM: one row of it at a time in LDS instead of a*a/8 bytes.
//  * w below thr: the sorted-list probes of tc.cu:273-286.  For a vertex v
//    with k entries of adj+(v) below thr these are the first
//    C(d,2) - C(d-k,2) wedges of its row-major order, so the list part is a
//    wedge order of its own -- a prefix of every vertex's -- with its own
//    prefix (the GraphDesc's ws), counted by the row walk of k_walk.
//
// A wedge range [w_begin, w_end) of the full order (the reference's wedge
// numbering, tc.cu:203-212) maps to one contiguous range of the list part's
// order and to the in-edges of the vertices holding it, with the tails of
// the two end vertices clipped.
//
// Pipeline on one stream: k_split_partials, k_scan_partials (three times)
// and k_split_scan (the two wedge prefixes and the top in-edge prefix);
// k_split_range; k_setup (tiles and search tree of the list part); the
// transposition in two counting-sort levels, k_split_l1<count>,
// k_split_l1scan, k_split_l1<scatter> (records bucketed by part of
// kSplitPart rows), k_split_l2<count>, k_split_l2cols, k_split_rowscan,
// k_split_l2<scatter> (each part's records grouped by row); k_split_rows
// (top-a part); k_walk (list part).
#include "wtc_dev.h"
#include "wtc_kernels.h"
#include "wtc_launch.h"

namespace wtc {

namespace {

// Entries of the ascending row colInd[b, e) below x.
template <typename OffT>
__device__ __forceinline__ uint32_t entries_below(const uint32_t *__restrict__ ci, OffT b, OffT e, uint32_t x) {
    if (b == e || ci[b] >= x) return 0u;
    if (ci[e - 1] < x) return (uint32_t)(e - b);
    OffT lo = b + 1, hi = e - 1;  // the first entry >= x lies in [lo, hi]
    while (lo < hi) {
        const OffT mid = lo + ((hi - lo) >> 1);
        if (ci[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return (uint32_t)(lo - b);
}

// List-part wedges of a vertex with out-degree d and k entries below thr:
// the pairs (i, t), i < t < d, with i < k.
__device__ __forceinline__ uint64_t list_wedges(uint64_t d, uint64_t k) { return binom2_d(d) - binom2_d(d - k); }

// Position in the list part's order of position x of the full order.
__device__ uint64_t list_pos(const uint64_t *ws, const uint64_t *wf, uint32_t n, uint64_t x) {
    if (x >= wf[n]) return ws[n];
    const uint32_t v = rightmost_le(wf, 0u, n - 1, x);
    const uint64_t off = x - wf[v], cnt = ws[v + 1] - ws[v];
    return ws[v] + (off < cnt ? off : cnt);
}

// Four ids loaded by one 16-byte load.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <typename OffT>
struct TailPair {
    OffT lo, hi;  // colInd positions [lo, hi) of one in-edge's tail
};

}  // namespace

// --------------------------------------------------------------- prefixes

// Top in-edges of a row of out-degree d with k entries below thr: the
// positions j >= k of its entries that are not the row's last.
__device__ __forceinline__ uint64_t top_in_edges(uint64_t d, uint64_t k) { return d > k + 1 ? d - k - 1 : 0; }

// Block b: vertices [b*kScanChunk, (b+1)*kScanChunk): kpos, and the sums of
// the list part's wedges, the full order's wedges and the top in-edges.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_split_partials(const GraphDesc *__restrict__ desc,
                                                           const SplitDesc *__restrict__ sd,
                                                           uint64_t *__restrict__ part_list,
                                                           uint64_t *__restrict__ part_full,
                                                           uint64_t *__restrict__ part_in) {
    __shared__ uint64_t s_l[kWavesPerBlock], s_f[kWavesPerBlock], s_i[kWavesPerBlock];
    const GraphDesc &D = desc[0];
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(D.rowPtr);
    const uint32_t *__restrict__ colInd = D.colInd;
    uint32_t *__restrict__ kpos = sd->kpos;
    const uint32_t n = D.n, thr = D.thr;
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    uint64_t al = 0, af = 0, ai = 0;
    for (int k = 0; k < kScanItems; k++) {
        const uint64_t v = base + (uint64_t)k * kBlock + threadIdx.x;
        if (v < n) {
            const OffT b = rowPtr[v], e = rowPtr[v + 1];
            const uint32_t kb = entries_below<OffT>(colInd, b, e, thr);
            kpos[v] = kb;
            const uint64_t d = (uint64_t)(e - b);
            al += list_wedges(d, kb);
            af += binom2_d(d);
            ai += top_in_edges(d, kb);
        }
    }
    al = wave_sum(al);
    af = wave_sum(af);
    ai = wave_sum(ai);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        s_l[threadIdx.x >> 6] = al;
        s_f[threadIdx.x >> 6] = af;
        s_i[threadIdx.x >> 6] = ai;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part_list[blockIdx.x] = s_l[0] + s_l[1] + s_l[2] + s_l[3];
        part_full[blockIdx.x] = s_f[0] + s_f[1] + s_f[2] + s_f[3];
        part_in[blockIdx.x] = s_i[0] + s_i[1] + s_i[2] + s_i[3];
    }
}

// ws (list part), ws_full and ie_pre from the scanned block partials.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_split_scan(const GraphDesc *__restrict__ desc,
                                                       const SplitDesc *__restrict__ sd,
                                                       const uint64_t *__restrict__ part_list,
                                                       const uint64_t *__restrict__ part_full,
                                                       const uint64_t *__restrict__ part_in) {
    __shared__ uint64_t s_tot[kWavesPerBlock];
    const GraphDesc &D = desc[0];
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(D.rowPtr);
    const uint32_t *__restrict__ kpos = sd->kpos;
    uint64_t *__restrict__ ws = D.ws;
    uint64_t *__restrict__ wf = sd->ws_full;
    uint64_t *__restrict__ wi = sd->ie_pre;
    const uint32_t n = D.n;
    uint64_t cl = part_list[blockIdx.x], cf = part_full[blockIdx.x], ci = part_in[blockIdx.x];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    for (int k = 0; k < kScanItems; k++) {
        if (base + (uint64_t)k * kBlock >= n) break;  // block-uniform
        const uint64_t v = base + (uint64_t)k * kBlock + threadIdx.x;
        uint64_t xl = 0, xf = 0, xi = 0;
        if (v < n) {
            const uint64_t d = (uint64_t)(rowPtr[v + 1] - rowPtr[v]);
            const uint32_t kb = kpos[v];
            xl = list_wedges(d, kb);
            xf = binom2_d(d);
            xi = top_in_edges(d, kb);
        }
        uint64_t tl, tf, ti;
        const uint64_t el = block_excl_scan(xl, s_tot, tl);
        const uint64_t ef = block_excl_scan(xf, s_tot, tf);
        const uint64_t ei = block_excl_scan(xi, s_tot, ti);
        if (v < n) {
            ws[v] = cl + el;
            wf[v] = cf + ef;
            wi[v] = ci + ei;
        }
        cl += tl;
        cf += tf;
        ci += ti;
    }
    if (blockIdx.x == (n - 1) / kScanChunk && threadIdx.x == 0) {
        ws[n] = cl;
        wf[n] = cf;
        wi[n] = ci;
    }
}

// One thread: the counted range of the full order, its vertices and its
// image in the list part's order (which k_setup then tiles).
__global__ void k_split_range(GraphDesc *__restrict__ desc, SplitDesc *__restrict__ sd) {
    if (threadIdx.x != 0) return;
    GraphDesc &D = desc[0];
    SplitDesc &S = sd[0];
    const uint32_t n = D.n;
    const uint64_t *ws = D.ws, *wf = S.ws_full;
    const uint64_t W = wf[n];
    const uint64_t lo = D.w_begin < W ? D.w_begin : W, hi = D.w_end < W ? D.w_end : W;
    S.w_lo = lo;
    S.w_hi = hi;
    if (lo == 0 && hi == W) {  // the whole order: nothing clipped
        S.clip = 0;
        S.v_lo = 0;
        S.v_hi = n - 1;
        D.w_begin = 0;
        D.w_end = ws[n];
        return;
    }
    S.clip = 1;
    if (lo >= hi) {  // nothing to count
        S.v_lo = 1;
        S.v_hi = 0;
        D.w_begin = 0;
        D.w_end = 0;
        return;
    }
    S.v_lo = rightmost_le(wf, 0u, n - 1, lo);
    S.v_hi = rightmost_le(wf, 0u, n - 1, hi - 1);
    D.w_begin = list_pos(ws, wf, n, lo);
    D.w_end = list_pos(ws, wf, n, hi);
}

// ------------------------------------------------------------ transposition

// First vertex of chunk c of `chunks`: the vertex holding top in-edge
// floor(c*K/chunks) (K = all top in-edges), so that chunks hold about equal
// shares of the in-edges; chunk `chunks` starts at n.  Wave-uniform.
__device__ __forceinline__ uint32_t chunk_start(cptr<uint64_t> ie_pre, uint32_t n, uint32_t c, uint32_t chunks) {
    if (c >= chunks) return n;
    const uint64_t K = ie_pre[n];
    const uint64_t t = (uint64_t)(((unsigned __int128)K * c) / chunks);
    return rightmost_le(ie_pre, 0u, n - 1, t);
}

// Clips the tail [j+1, e) of in-edge j of an end vertex of the counted
// range (row start rb, full-order base ws) to its wedges in [w_lo, w_hi):
// the tail ids of row i = j - rb are the wedges [base, base + L) of the full
// order (tc.cu:203-212 numbering).  Returns the clipped [lo, hi) (lo == hi:
// none).
template <typename OffT>
__device__ __forceinline__ void clip_tail(OffT j, OffT e, OffT rb, uint64_t ws, uint64_t w_lo, uint64_t w_hi,
                                          OffT &lo, OffT &hi) {
    const uint64_t d = (uint64_t)(e - rb), i = (uint64_t)(j - rb);
    const uint64_t bse = ws + i * (d - 1) - binom2_d(i), L = (uint64_t)(e - j - 1);
    const uint64_t a = bse > w_lo ? bse : w_lo;
    const uint64_t z = bse + L < w_hi ? bse + L : w_hi;
    lo = j + 1;
    hi = lo;
    if (a < z) {
        lo = j + 1 + (OffT)(a - bse);
        hi = j + 1 + (OffT)(z - bse);
    }
}

// Vertices with at least this many top in-edges are taken by a whole wave.
constexpr uint32_t kSplitBigVertex = kWave;

// First level of the transposition.  Block b takes vertex chunk b (clipped
// to the counted vertices); a wave takes 64 consecutive vertices:
//  * a vertex with kSplitBigVertex or more top in-edges is taken by the
//    whole wave in turn: the lanes load its in-edges' ids 64 consecutive
//    positions at a time, four loads in flight;
//  * the others' in-edges are scanned and dealt to the lanes 64 at a time
//    (a lane finds its vertex by a binary lifting over the scanned counts).
// Every in-edge (its tail clipped at the ends of the counted range; empty
// ones dropped) belongs to the part (id - thr) / kSplitPart.  kScatter =
// false: the block counts its in-edges per part into hist1[b][part];
// kScatter = true (hist1 then holding the record offset of each chunk and
// part): each in-edge takes the next record slot of its part through an LDS
// cursor, and its id and tail are stored there.  Both passes see the same
// in-edges in the same chunks.
template <typename OffT, bool kScatter>
__global__ __launch_bounds__(kSplitBlock) void k_split_l1(const GraphDesc *__restrict__ desc,
                                                          const SplitDesc *__restrict__ sd) {
    __shared__ uint32_t s_cnt[kSplitPartsMax];
    __shared__ uint64_t s_base[kSplitPartsMax];
    const GraphDesc &D = desc[0];
    const SplitDesc &S = sd[0];
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(D.rowPtr);
    const gptr<uint32_t> colInd = to_global(D.colInd);
    const uint32_t *__restrict__ kpos = S.kpos;
    const uint64_t *__restrict__ wf = S.ws_full;
    uint64_t *__restrict__ hist1 = S.hist1;
    uint32_t *__restrict__ rec_id = S.rec_id;
    TailPair<OffT> *__restrict__ rec = static_cast<TailPair<OffT> *>(S.rec);
    const uint32_t n = D.n, thr = D.thr, parts = S.parts;
    const uint32_t b = blockIdx.x;
    for (uint32_t i = threadIdx.x; i < parts; i += kSplitBlock) {
        s_cnt[i] = 0u;
        s_base[i] = kScatter ? hist1[b * parts + i] : 0u;
    }
    __syncthreads();

    // one in-edge: count it in its part, or store its record
    auto take = [&](uint32_t id, OffT lo, OffT hi) {
        const uint32_t q = (id - thr) / kSplitPart;
        if (!kScatter) {
            atomicAdd(&s_cnt[q], 1u);
        } else {
            const uint64_t p = s_base[q] + atomicAdd(&s_cnt[q], 1u);
            rec_id[p] = id;
            rec[p] = TailPair<OffT>{lo, hi};
        }
    };

    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x >> 6;
    const bool clip = S.clip != 0;
    const uint32_t v_lo = S.v_lo, v_hi = S.v_hi;
    const uint64_t w_lo = S.w_lo, w_hi = S.w_hi;
    const cptr<uint64_t> ie_pre = to_const(static_cast<const uint64_t *>(S.ie_pre));
    uint64_t c0 = uni32(chunk_start(ie_pre, n, b, S.chunks)), c1 = uni32(chunk_start(ie_pre, n, b + 1, S.chunks));
    if (clip) {
        c0 = c0 > v_lo ? c0 : v_lo;
        c1 = c1 < (uint64_t)v_hi + 1 ? c1 : (uint64_t)v_hi + 1;
    }
    constexpr uint32_t kWaves = kSplitBlock / kWave;
    for (uint64_t g = c0 + (uint64_t)wv * kWave; g < c1; g += (uint64_t)kWaves * kWave) {
        const uint64_t v = g + lane;
        OffT j0 = 0, e = 0, rb = 0;
        uint32_t cnt = 0;
        bool edge = false;
        uint64_t wsv = 0;
        if (v < c1) {
            rb = rowPtr[v];
            e = rowPtr[v + 1];
            j0 = rb + kpos[v];  // the row's first id >= thr
            cnt = (uint32_t)top_in_edges((uint64_t)(e - rb), kpos[v]);
            edge = clip && (v == v_lo || v == v_hi);
            if (edge) wsv = wf[v];
        }
        // big vertices, one at a time by the whole wave
        uint64_t big = __ballot(cnt >= kSplitBigVertex);
        while (big) {
            const uint32_t src = (uint32_t)__builtin_ctzll(big);
            big &= big - 1;
            const OffT vj0 = (OffT)readlane_u64((uint64_t)j0, src);
            const OffT vend = vj0 + (OffT)(uint32_t)__builtin_amdgcn_readlane((int)cnt, (int)src);
            const OffT ve = (OffT)readlane_u64((uint64_t)e, src);
            const bool vedge = __builtin_amdgcn_readlane((int)edge, (int)src) != 0;
            const OffT vrb = (OffT)readlane_u64((uint64_t)rb, src);
            const uint64_t vws = readlane_u64(wsv, src);
            for (OffT jb = vj0; jb < vend; jb += 4 * kWave) {
                uint32_t ids[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const OffT j = jb + (OffT)(u * kWave) + lane;
                    ids[u] = j < vend ? colInd[j] : 0u;
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const OffT j = jb + (OffT)(u * kWave) + lane;
                    if (j < vend) {
                        OffT lo = j + 1, hi = ve;
                        if (vedge) clip_tail<OffT>(j, ve, vrb, vws, w_lo, w_hi, lo, hi);
                        if (lo < hi) take(ids[u], lo, hi);
                    }
                }
            }
        }
        if (cnt >= kSplitBigVertex) cnt = 0;
        // the other vertices' in-edges, dealt to the lanes
        uint32_t incl = cnt;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)incl, o, kWave);
            if ((int)lane >= o) incl += y;
        }
        const uint32_t tot = uni32((uint32_t)__shfl((int)incl, kWave - 1, kWave));
        const uint32_t excl = incl - cnt;
        const bool any_edge = __ballot(edge) != 0;  // wave-uniform
        for (uint32_t base = 0; base < tot; base += kWave) {
            const uint32_t idx = base + lane;
            // the lane's vertex: the rightmost whose first in-edge is <= idx
            uint32_t pos = 0;
#pragma unroll
            for (uint32_t st = kWave / 2; st > 0; st >>= 1) {
                const uint32_t c = pos + st;
                const uint32_t x = (uint32_t)__shfl((int)excl, (int)(c & (kWave - 1)), kWave);
                if (c < (uint32_t)kWave && x <= idx) pos = c;
            }
            const OffT j = shfl_off(j0, pos) + (OffT)(idx - (uint32_t)__shfl((int)excl, (int)pos, kWave));
            const OffT ev = shfl_off(e, pos);
            OffT lo = j + 1, hi = ev;
            if (any_edge) {
                const bool ed = __shfl((int)edge, (int)pos, kWave) != 0;
                const OffT rbv = shfl_off(rb, pos);
                const uint64_t wv0 = shfl_u64(wsv, pos);
                if (ed) clip_tail<OffT>(j, ev, rbv, wv0, w_lo, w_hi, lo, hi);
            }
            if (idx < tot && lo < hi) take(colInd[j], lo, hi);
        }
    }
    if (!kScatter) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < parts; i += kSplitBlock) hist1[b * parts + i] = s_cnt[i];
    }
}

// One thread per part (one block): the records of each part (its column of
// hist1 summed), the parts' first records (part_base), each chunk's record
// offset in each part (hist1 rewritten), and the first second-level block
// of each part (c2base: kSplitChunk2 records per block).
__global__ __launch_bounds__(kSplitPartsMax) void k_split_l1scan(const SplitDesc *__restrict__ sd) {
    __shared__ uint64_t s_tot[kSplitPartsMax];
    const SplitDesc &S = sd[0];
    const uint32_t parts = S.parts, chunks = S.chunks, q = threadIdx.x;
    uint64_t *__restrict__ hist1 = S.hist1;
    uint64_t tot = 0;
    if (q < parts)
        for (uint32_t c = 0; c < chunks; c++) tot += hist1[c * parts + q];
    s_tot[q] = q < parts ? tot : 0;
    __syncthreads();
    if (q == 0) {
        uint64_t base = 0;
        uint32_t c2 = 0;
        for (uint32_t i = 0; i < parts; i++) {
            const uint64_t t = s_tot[i];
            S.part_base[i] = base;
            S.c2base[i] = c2;
            s_tot[i] = base;
            base += t;
            c2 += (uint32_t)((t + kSplitChunk2 - 1) / kSplitChunk2);
        }
        S.part_base[parts] = base;
        S.c2base[parts] = c2;
    }
    __syncthreads();
    if (q < parts) {
        uint64_t off = s_tot[q];
        for (uint32_t c = 0; c < chunks; c++) {
            const uint64_t h = hist1[c * parts + q];
            hist1[c * parts + q] = off;
            off += h;
        }
    }
}

// Second level.  Block c2 takes records [part_base[q] + k*kSplitChunk2, ...)
// of part q (c2 = c2base[q] + k; blocks past c2base[parts] do nothing) and
// groups them by row of the part: kScatter = false counts them per row into
// hist2[c2][row - q*kSplitPart]; kScatter = true (hist2 then holding, per
// row, the part's records of earlier blocks) places each record's tail at
// the next slot of its row in ie.
template <typename OffT, bool kScatter>
__global__ __launch_bounds__(kSplitBlock) void k_split_l2(const GraphDesc *__restrict__ desc,
                                                          const SplitDesc *__restrict__ sd) {
    __shared__ uint32_t s_cnt[kSplitPart];
    const SplitDesc &S = sd[0];
    const uint32_t parts = S.parts, rows = S.rows, thr = desc[0].thr;
    const uint32_t c2 = blockIdx.x;
    if (c2 >= S.c2base[parts]) return;  // block-uniform
    uint32_t q = 0;
    while (q + 1 < parts && S.c2base[q + 1] <= c2) q++;
    const uint32_t r0 = q * kSplitPart, r1 = r0 + kSplitPart < rows ? r0 + kSplitPart : rows;
    const uint64_t k0 = S.part_base[q] + (uint64_t)(c2 - S.c2base[q]) * kSplitChunk2;
    const uint64_t kq = S.part_base[q + 1], k1 = k0 + kSplitChunk2 < kq ? k0 + kSplitChunk2 : kq;
    uint32_t *__restrict__ hist2 = S.hist2 + (uint64_t)c2 * kSplitPart;
    for (uint32_t i = threadIdx.x; i < r1 - r0; i += kSplitBlock) s_cnt[i] = kScatter ? hist2[i] : 0u;
    __syncthreads();
    const uint32_t *__restrict__ rec_id = S.rec_id;
    const TailPair<OffT> *__restrict__ rec = static_cast<const TailPair<OffT> *>(S.rec);
    const OffT *__restrict__ rp_in = static_cast<const OffT *>(S.rp_in);
    TailPair<OffT> *__restrict__ ie = static_cast<TailPair<OffT> *>(S.ie);
    for (uint64_t k = k0 + threadIdx.x; k < k1; k += kSplitBlock) {
        const uint32_t r = rec_id[k] - thr;
        if (!kScatter) {
            atomicAdd(&s_cnt[r - r0], 1u);
        } else {
            const OffT p = rp_in[r] + (OffT)atomicAdd(&s_cnt[r - r0], 1u);
            ie[p] = rec[k];
        }
    }
    if (!kScatter) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < r1 - r0; i += kSplitBlock) hist2[i] = s_cnt[i];
    }
}

// Thread per top row r (part q): the column prefix of hist2 over the part's
// second-level blocks (hist2[c2][r - q*kSplitPart] becomes the row's records
// in earlier blocks) and the row's total.
__global__ __launch_bounds__(kBlock) void k_split_l2cols(const SplitDesc *__restrict__ sd) {
    const SplitDesc &S = sd[0];
    const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= S.rows) return;
    const uint32_t q = r / kSplitPart, i = r - q * kSplitPart;
    uint32_t *__restrict__ hist2 = S.hist2;
    uint32_t s = 0;
    for (uint32_t c2 = S.c2base[q]; c2 < S.c2base[q + 1]; c2++) {
        const uint32_t h = hist2[(uint64_t)c2 * kSplitPart + i];
        hist2[(uint64_t)c2 * kSplitPart + i] = s;
        s += h;
    }
    S.row_tot[r] = s;
}

// One block: rp_in = exclusive prefix of the row totals; rp_in[rows] = the
// in-edges; the row kernel's tile counter zeroed; the counted wedges of the
// full order, those of the top rows and the in-edges into meta.
template <typename OffT>
__global__ __launch_bounds__(kSplitBlock) void k_split_rowscan(const GraphDesc *__restrict__ desc,
                                                               const SplitDesc *__restrict__ sd,
                                                               uint64_t *__restrict__ meta) {
    constexpr int kWaves = kSplitBlock / kWave;
    __shared__ uint64_t s_w[kWaves];
    const SplitDesc &S = sd[0];
    const uint32_t rows = S.rows;
    const uint32_t *__restrict__ tot = S.row_tot;
    OffT *__restrict__ rp = static_cast<OffT *>(S.rp_in);
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x >> 6;
    uint64_t carry = 0;
    for (uint32_t base = 0; base < rows; base += kSplitBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint64_t x = i < rows ? tot[i] : 0;
        const uint64_t inc = wave_incl_scan(x);
        if (lane == kWave - 1) s_w[wv] = inc;
        __syncthreads();
        uint64_t off = 0, all = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) {
            const uint64_t t = s_w[w];
            off += w < (int)wv ? t : 0;
            all += t;
        }
        __syncthreads();
        if (i < rows) rp[i] = (OffT)(carry + off + inc - x);
        carry += all;
    }
    if (threadIdx.x == 0) {
        rp[rows] = (OffT)carry;
        meta[kMetaRows] = 0;
        meta[kMetaWedges] = S.w_hi - S.w_lo;  // k_setup counted the list part's only
        meta[kMetaTopWedges] = S.w_hi - S.w_lo - desc[0].wcount;
        meta[kMetaInEdges] = carry;
    }
}

// ------------------------------------------------------------- top-a part

// Lanes per in-edge tail in the row kernel: a wave takes kWave / kSplitLanes
// in-edges per step.  Tails average ~180 ids on R-MAT-24; with a whole wave
// per tail most lanes idled (37.4 ms vs 25.1 ms at 16 lanes, 28.1 at 32,
// 24.8 at 8).
constexpr int kSplitLanes = 16;

// Persistent blocks take tiles of kSplitTile consecutive in-edges (in row
// order) from one counter.  Per row of a tile: adj+(w) set as bits (u - thr)
// in LDS, the tile's in-edges of the row dealt to the waves (kLanes lanes
// per in-edge), and each tail streamed through the bitmap by 16-byte loads,
// two in flight per lane; then the row's bits cleared.  colInd must be
// 16-byte aligned and readable 12 bytes past its last id (every buffer of
// this library carries 256 bytes of slack).

template <typename OffT, int kWaves, int kLanes>
__global__ __launch_bounds__(kWave * kWaves) void k_split_rows(const GraphDesc *__restrict__ desc,
                                                               const SplitDesc *__restrict__ sd,
                                                               uint64_t *__restrict__ meta,
                                                               unsigned long long *__restrict__ counts) {
    extern __shared__ uint32_t s_bits[];
    __shared__ uint64_t s_tile;
    const GraphDesc &D = desc[0];
    const SplitDesc &S = sd[0];
    const cptr<OffT> rowPtrU = to_const(static_cast<const OffT *>(D.rowPtr));
    const gptr<uint32_t> colInd = to_global(D.colInd);
    const cptr<OffT> rp_in = to_const(static_cast<const OffT *>(S.rp_in));
    const cptr<OffT> ieU = to_const(static_cast<const OffT *>(S.ie));
    const gptr<OffT> ieG = to_global(static_cast<const OffT *>(S.ie));
    const uint32_t thr = D.thr, rows = S.rows;
    constexpr uint32_t kPer = kWave / kLanes;  // in-edges a wave takes per step
    const uint32_t words = (rows + 31) >> 5;
    for (uint32_t i = threadIdx.x; i < words; i += kWave * kWaves) s_bits[i] = 0;
    const uint64_t K = (uint64_t)rp_in[rows];
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = uni32(threadIdx.x >> 6);
    const uint32_t sub = lane / kLanes, sl = lane % kLanes;
    unsigned long long *const ctr = reinterpret_cast<unsigned long long *>(meta + kMetaRows);
    uint64_t hits = 0;
    for (;;) {
        __syncthreads();  // the previous tile's s_tile reads and bit clears are done
        if (threadIdx.x == 0) s_tile = atomicAdd(ctr, 1ull);
        __syncthreads();
        const uint64_t k0 = uni64(s_tile) * kSplitTile;
        if (k0 >= K) break;  // block-uniform
        const uint64_t k1 = k0 + kSplitTile < K ? k0 + kSplitTile : K;
        uint64_t k = k0;
        while (k < k1) {
            const uint32_t r = uni32(rightmost_le(rp_in, 0u, rows - 1, k));  // the row holding in-edge k
            const uint64_t re = (uint64_t)rp_in[r + 1];
            const uint64_t kend = re < k1 ? re : k1;
            const uint32_t w = thr + r;
            const OffT ab = rowPtrU[w], ae = rowPtrU[w + 1];
            for (OffT i = ab + threadIdx.x; i < ae; i += kWave * kWaves) {
                const uint32_t c = colInd[i] - thr;
                atomicOr(&s_bits[c >> 5], 1u << (c & 31u));
            }
            __syncthreads();
            uint32_t h = 0;
            for (uint64_t x0 = k + (uint64_t)wv * kPer; x0 < kend; x0 += (uint64_t)kWaves * kPer) {
                // kPer in-edges per step, kLanes lanes each
                OffT lo = 0, hi = 0;
                if (kPer == 1) {
                    lo = uni(ieU[2 * x0]);
                    hi = uni(ieU[2 * x0 + 1]);
                } else if (x0 + sub < kend) {
                    lo = ieG[2 * (x0 + sub)];
                    hi = ieG[2 * (x0 + sub) + 1];
                }
                for (OffT q = (lo & ~(OffT)3) + 4 * sl; q < hi; q += 8 * kLanes) {
                    const OffT q1 = q + 4 * kLanes;
                    const u32x4 a = *reinterpret_cast<gptr<u32x4>>(colInd + q);
                    u32x4 b = {0u, 0u, 0u, 0u};
                    if (q1 < hi) b = *reinterpret_cast<gptr<u32x4>>(colInd + q1);
                    const uint32_t ua[8] = {a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
#pragma unroll
                    for (int z = 0; z < 8; z++) {
                        const OffT t = (z < 4 ? q : q1) + (OffT)(z & 3);
                        if (t >= lo && t < hi) {
                            const uint32_t c = ua[z] - thr;
                            h += (s_bits[c >> 5] >> (c & 31u)) & 1u;
                        }
                    }
                }
            }
            hits += h;
            __syncthreads();
            for (OffT i = ab + threadIdx.x; i < ae; i += kWave * kWaves) s_bits[(colInd[i] - thr) >> 5] = 0u;
            __syncthreads();
            k = kend;
        }
    }
    hits = wave_sum(hits);
    if (lane == 0 && hits) atomicAdd(&counts[0], (unsigned long long)hits);
}

// ---------------------------------------------------------------- launches

hipError_t launch_split_prefix(const GraphDesc *desc, const SplitDesc *sd, int ob, uint32_t nblocks,
                               uint64_t *partials, hipStream_t st) {
    if (nblocks == 0) return hipSuccess;
    uint64_t *pl = partials, *pf = partials + nblocks, *pi = partials + 2 * (size_t)nblocks;
    hipError_t e = WTC_LAUNCH_OFF(ob, k_split_partials, dim3(nblocks), dim3(kBlock), 0, st, desc, sd, pl, pf, pi);
    if (e != hipSuccess) return e;
    if ((e = launch_scan_partials(pl, nblocks, st)) != hipSuccess) return e;
    if ((e = launch_scan_partials(pf, nblocks, st)) != hipSuccess) return e;
    if ((e = launch_scan_partials(pi, nblocks, st)) != hipSuccess) return e;
    return WTC_LAUNCH_OFF(ob, k_split_scan, dim3(nblocks), dim3(kBlock), 0, st, desc, sd, (const uint64_t *)pl,
                          (const uint64_t *)pf, (const uint64_t *)pi);
}

hipError_t launch_split_range(GraphDesc *desc, SplitDesc *sd, hipStream_t st) {
    k_split_range<<<dim3(1), dim3(kWave), 0, st>>>(desc, sd);
    return hipGetLastError();
}

hipError_t launch_split_transpose(const GraphDesc *desc, const SplitDesc *sd, int ob, uint32_t chunks,
                                  uint32_t c2cap, uint32_t rows, uint64_t *meta, hipStream_t st) {
    if (rows == 0 || chunks == 0 || c2cap == 0) return hipErrorInvalidValue;
    const dim3 blk(kSplitBlock), g1(chunks), g2(c2cap);
    hipError_t e;
    if (ob == 8) k_split_l1<uint64_t, false><<<g1, blk, 0, st>>>(desc, sd);
    else k_split_l1<uint32_t, false><<<g1, blk, 0, st>>>(desc, sd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_split_l1scan<<<dim3(1), dim3(kSplitPartsMax), 0, st>>>(sd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ob == 8) k_split_l1<uint64_t, true><<<g1, blk, 0, st>>>(desc, sd);
    else k_split_l1<uint32_t, true><<<g1, blk, 0, st>>>(desc, sd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ob == 8) k_split_l2<uint64_t, false><<<g2, blk, 0, st>>>(desc, sd);
    else k_split_l2<uint32_t, false><<<g2, blk, 0, st>>>(desc, sd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_split_l2cols<<<dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0, st>>>(sd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ob == 8) k_split_rowscan<uint64_t><<<dim3(1), blk, 0, st>>>(desc, sd, meta);
    else k_split_rowscan<uint32_t><<<dim3(1), blk, 0, st>>>(desc, sd, meta);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ob == 8) k_split_l2<uint64_t, true><<<g2, blk, 0, st>>>(desc, sd);
    else k_split_l2<uint32_t, true><<<g2, blk, 0, st>>>(desc, sd);
    return hipGetLastError();
}

size_t split_rows_lds(uint32_t rows) { return (size_t)((rows + 31) / 32) * sizeof(uint32_t); }

template <int W>
static const void *split_rows_fn(int ob) {
    return ob == 8 ? reinterpret_cast<const void *>(&k_split_rows<uint64_t, W, kSplitLanes>)
                   : reinterpret_cast<const void *>(&k_split_rows<uint32_t, W, kSplitLanes>);
}

static const void *split_rows_kernel(int ob, int wpb) {
    switch (wpb) {
        case 4: return split_rows_fn<4>(ob);
        case 16: return split_rows_fn<16>(ob);
        default: return split_rows_fn<8>(ob);
    }
}

bool valid_split_waves(int wpb) { return wpb == 4 || wpb == 8 || wpb == 16; }

hipError_t split_rows_grid(int ob, uint32_t rows, int wpb, int num_cus, uint32_t *grid) {
    if (!valid_split_waves(wpb)) return hipErrorInvalidValue;
    const void *fn = split_rows_kernel(ob, wpb);
    const size_t shm = split_rows_lds(rows);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, wpb * kWave, shm);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;  // the row bitmap does not fit LDS
    *grid = (uint32_t)per_cu * (uint32_t)(num_cus > 0 ? num_cus : 1);
    return hipSuccess;
}

hipError_t launch_split_rows(const GraphDesc *desc, const SplitDesc *sd, int ob, uint32_t rows, int wpb,
                             uint64_t *meta, unsigned long long *counts, uint32_t grid, hipStream_t st) {
    if (rows == 0) return hipSuccess;
    if (!valid_split_waves(wpb) || grid == 0) return hipErrorInvalidValue;
    const void *fn = split_rows_kernel(ob, wpb);
    void *args[] = {&desc, &sd, &meta, &counts};
    return hipLaunchKernel(fn, dim3(grid), dim3(wpb * kWave), args, split_rows_lds(rows), st);
}

}  // namespace wtc
