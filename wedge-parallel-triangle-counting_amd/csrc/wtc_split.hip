// wtc_split.hip -- the split count of one graph (counting kernel choice
// kKernelSplit).  Every wedge (v, w, u) of a degree-oriented CSR belongs to
// one in-edge v -> w: w = colInd[j] at a position j of adj+(v) that is not
// the row's last, and u runs over the row's tail colInd[j+1 .. e).  The
// in-edge closes |tail ∩ adj+(w)| triangles.  The in-edges are grouped by
// their middle vertex w with a counting sort built per count, and counted in
// two kernels chosen by w:
//
//  * w among the top-a ranks (w >= thr): the wedges whose probe the
//    reference answers from its top-a bitmap (tc.cu:262-272).  One key per
//    row.  A block of the row kernel holds adj+(w) as a bitmap of a bits in
//    LDS, and its waves stream the tails of w's in-edges through it from a
//    packed copy of colInd (ids above thr as 21-bit offsets, six per 16-byte
//    load).  Each probe is one LDS read, and the top-a bitmap never exists
//    in HBM: one row of it at a time in LDS instead of a*a/8 bytes.
//  * w below thr: the sorted-list probes of tc.cu:273-286.  One key per slab
//    of colInd (kSlabSpan positions; a row belongs to the slab its adjacency
//    starts in).  A block of the slab kernel hashes the rows of the slab its
//    in-edges name into LDS, and its waves stream the tails of the slab's
//    in-edges, looking each id up in adj+(w)'s hash set.
//
// A wedge range [w_begin, w_end) of the full order (the reference's wedge
// numbering, tc.cu:203-212) holds the in-edges of the vertices between its
// ends, with the tails of the two end vertices clipped.
//
// Pipeline on one stream: k_split_partials, k_scan_partials (twice) and
// k_split_scan (the full wedge prefix and the in-edge prefix);
// k_split_range; the transposition in three counting-sort levels,
// k_split_l1count, k_split_l1cols, k_split_l1bases, k_split_l1
// (records bucketed by part of kSplitPart keys), k_split_level_count<2>,
// k_split_levelcols<2>, k_split_subscan, k_split_level_scatter<2> (each
// part's records bucketed by sub-part of kSplitSub keys),
// k_split_level_count<3>, k_split_levelcols<3>, k_split_keysum,
// k_scan_partials, k_split_keyscan, k_split_level_scatter<3> (each
// sub-part's records grouped by key), k_split_pack (colInd as 21-bit ids
// above thr, over the second level's records);
// k_split_rows (top rows); k_split_slabs (list rows).
#include <type_traits>
#include <utility>

#include "wtc_dev.h"
#include "wtc_kernels.h"
#include "wtc_launch.h"

namespace wtc {

namespace {

// Four ids loaded by one 16-byte load.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <typename OffT>
struct TailPair {
    OffT lo, hi;  // a top record: colInd positions [lo, hi) of one in-edge's tail;
                  // a list record: (j, hi), the tail [j + 1, hi)
};

// Wedges before row i of a vertex of out-degree d in its row-major order
// (row r holds d - 1 - r wedges).
__device__ __forceinline__ uint64_t row_base(uint64_t i, uint64_t d) { return i * (d - 1) - binom2_d(i); }

}  // namespace

// --------------------------------------------------------------- prefixes

// In-edges of a row of out-degree d: every position but the last.
__device__ __forceinline__ uint64_t in_edges_of(uint64_t d) { return d ? d - 1 : 0; }

// Block b: vertices [b*kScanChunk, (b+1)*kScanChunk): the sums of the full
// order's wedges and of the in-edges.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_split_partials(const GraphDesc *__restrict__ desc,
                                                           uint64_t *__restrict__ part_full,
                                                           uint64_t *__restrict__ part_in) {
    __shared__ uint64_t s_f[kWavesPerBlock], s_i[kWavesPerBlock];
    const GraphDesc &D = desc[0];
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(D.rowPtr);
    const uint32_t n = D.n;
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    uint64_t af = 0, ai = 0;
    for (int k = 0; k < kScanItems; k++) {
        const uint64_t v = base + (uint64_t)k * kBlock + threadIdx.x;
        if (v < n) {
            const uint64_t d = (uint64_t)(rowPtr[v + 1] - rowPtr[v]);
            af += binom2_d(d);
            ai += in_edges_of(d);
        }
    }
    af = wave_sum(af);
    ai = wave_sum(ai);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        s_f[threadIdx.x >> 6] = af;
        s_i[threadIdx.x >> 6] = ai;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part_full[blockIdx.x] = s_f[0] + s_f[1] + s_f[2] + s_f[3];
        part_in[blockIdx.x] = s_i[0] + s_i[1] + s_i[2] + s_i[3];
    }
}

// ws_full and ie_pre from the scanned block partials, and slab_row.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_split_scan(const GraphDesc *__restrict__ desc,
                                                       const SplitDesc *__restrict__ sd,
                                                       const uint64_t *__restrict__ part_full,
                                                       const uint64_t *__restrict__ part_in) {
    __shared__ uint64_t s_tot[kWavesPerBlock];
    const GraphDesc &D = desc[0];
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(D.rowPtr);
    uint64_t *__restrict__ wf = sd->ws_full;
    uint64_t *__restrict__ wi = sd->ie_pre;
    uint32_t *__restrict__ slab_row = sd->slab_row;
    const uint32_t n = D.n, thr = D.thr, slabs = sd->slabs;
    uint64_t cf = part_full[blockIdx.x], ci = part_in[blockIdx.x];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    for (int k = 0; k < kScanItems; k++) {
        if (base + (uint64_t)k * kBlock >= n) break;  // block-uniform
        const uint64_t v = base + (uint64_t)k * kBlock + threadIdx.x;
        uint64_t xf = 0, xi = 0;
        if (v < n) {
            const uint64_t d = (uint64_t)(rowPtr[v + 1] - rowPtr[v]);
            xf = binom2_d(d);
            xi = in_edges_of(d);
        }
        uint64_t tf, ti;
        const uint64_t ef = block_excl_scan(xf, s_tot, tf);
        const uint64_t ei = block_excl_scan(xi, s_tot, ti);
        if (v < n) {
            wf[v] = cf + ef;
            wi[v] = ci + ei;
            if (v <= thr) {  // the slabs whose first row is v: rowPtr[v - 1] < s*kSlabSpan <= rowPtr[v]
                const uint64_t s_hi = (uint64_t)rowPtr[v] / kSlabSpan;
                for (uint64_t s = v ? (uint64_t)rowPtr[v - 1] / kSlabSpan + 1 : 0; s <= s_hi && s <= slabs; s++)
                    slab_row[s] = (uint32_t)v;
            }
        }
        cf += tf;
        ci += ti;
    }
    if (blockIdx.x == (n - 1) / kScanChunk && threadIdx.x == 0) {
        wf[n] = cf;
        wi[n] = ci;
    }
}

// Thread 0: the counted range of the full order, its end vertices, and the
// one in-edge whose tail the range cuts at the front (row i0 of v_lo holding
// wedge w_lo: its tail starts w_lo - base(i0) ids in).  Thread q < parts:
// list_bid[q], the first list vertex w (below thr) whose key rows +
// rowPtr[w] / kSlabSpan lies in part q or later (keys grow with w).
template <typename OffT>
__global__ __launch_bounds__(kSplitPartsMax) void k_split_range(const GraphDesc *__restrict__ desc,
                                                                SplitDesc *__restrict__ sd) {
    const GraphDesc &D = desc[0];
    SplitDesc &S = sd[0];
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(D.rowPtr);
    const uint32_t q = threadIdx.x;
    if (q < S.parts) {
        const uint64_t first = S.rows / kSplitPart;  // the part of the first list key
        uint32_t bid = 0;
        if (q > first) {
            const uint64_t at = ((uint64_t)q * kSplitPart - S.rows) * kSlabSpan;  // colInd position
            uint32_t lo = 0, hi = D.thr;  // the first w in [0, thr] with rowPtr[w] >= at (thr: none below)
            while (lo < hi) {
                const uint32_t mid = lo + ((hi - lo) >> 1);
                if ((uint64_t)rowPtr[mid] < at) lo = mid + 1;
                else hi = mid;
            }
            bid = lo;
        }
        S.list_bid[q] = bid;
    }
    if (threadIdx.x != 0) return;
    const uint32_t n = D.n;
    const uint64_t *wf = S.ws_full;
    const uint64_t W = wf[n];
    const uint64_t lo = D.w_begin < W ? D.w_begin : W, hi = D.w_end < W ? D.w_end : W;
    S.w_lo = lo;
    S.w_hi = hi;
    S.clip_j = ~0ull;
    S.clip_lo = 0;
    if (lo == 0 && hi == W) {  // the whole order: nothing clipped
        S.clip = 0;
        S.v_lo = 0;
        S.v_hi = n - 1;
        return;
    }
    S.clip = 1;
    if (lo >= hi) {  // nothing to count
        S.v_lo = 1;
        S.v_hi = 0;
        return;
    }
    const uint32_t v = rightmost_le(wf, 0u, n - 1, lo);
    S.v_lo = v;
    S.v_hi = rightmost_le(wf, 0u, n - 1, hi - 1);
    const uint64_t rb = (uint64_t)rowPtr[v], d = (uint64_t)rowPtr[v + 1] - rb, off = lo - wf[v];
    uint64_t s = 0, e = d - 1;  // the last i in [0, d - 2] with row_base(i) <= off
    while (e - s > 1) {
        const uint64_t mid = s + ((e - s) >> 1);
        if (row_base(mid, d) <= off) s = mid;
        else e = mid;
    }
    S.clip_j = rb + s;
    S.clip_lo = rb + s + 1 + (off - row_base(s, d));
}

// ------------------------------------------------------------ transposition

// First vertex of chunk c of `chunks`: the vertex holding in-edge
// floor(c*K/chunks) (K = all in-edges), so that chunks hold about equal
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
    const uint64_t bse = ws + row_base(i, d), L = (uint64_t)(e - j - 1);
    const uint64_t a = bse > w_lo ? bse : w_lo;
    const uint64_t z = bse + L < w_hi ? bse + L : w_hi;
    lo = j + 1;
    hi = lo;
    if (a < z) {
        lo = j + 1 + (OffT)(a - bse);
        hi = j + 1 + (OffT)(z - bse);
    }
}

// First-level count, by positions: block b counts the in-edges of vertex
// chunk b (clipped to the counted vertices) per part into hist1[b][part],
// the same in-edges the scatter pass of k_split_l1 takes.  An in-edge is a
// position of adj+(v) other than the row's last, so the block counts every
// position of its chunk's rows (colInd read once, coalesced), takes back
// each row's last position, and, when the counted range clips its end
// vertices, takes back their in-edges whose clipped tail is empty.  A list
// vertex's part comes from list_bid (no rowPtr read).
template <typename OffT>
__global__ __launch_bounds__(kSplitBlock) void k_split_l1count(const GraphDesc *__restrict__ desc,
                                                               const SplitDesc *__restrict__ sd) {
    __shared__ uint32_t s_cnt[kSplitPartsMax];
    __shared__ uint32_t s_bid[kSplitPartsMax];
    const GraphDesc &D = desc[0];
    const SplitDesc &S = sd[0];
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(D.rowPtr);
    const gptr<uint32_t> colInd = to_global(D.colInd);
    const uint32_t n = D.n, thr = D.thr, parts = S.parts;
    const uint32_t b = blockIdx.x;
    for (uint32_t i = threadIdx.x; i < parts; i += kSplitBlock) {
        s_cnt[i] = 0u;
        s_bid[i] = S.list_bid[i];
    }
    __syncthreads();
    auto part_of = [&](uint32_t id) -> uint32_t {
        if (id >= thr) return (id - thr) / kSplitPart;
        uint32_t lo = 0, hi = parts;  // s_bid[lo] <= id < s_bid[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_bid[mid] <= id) lo = mid;
            else hi = mid;
        }
        return lo;
    };
    const bool clip = S.clip != 0;
    const uint32_t v_lo = S.v_lo, v_hi = S.v_hi;
    const cptr<uint64_t> ie_pre = to_const(static_cast<const uint64_t *>(S.ie_pre));
    uint64_t c0 = uni32(chunk_start(ie_pre, n, b, S.chunks)), c1 = uni32(chunk_start(ie_pre, n, b + 1, S.chunks));
    if (clip) {
        c0 = c0 > v_lo ? c0 : v_lo;
        c1 = c1 < (uint64_t)v_hi + 1 ? c1 : (uint64_t)v_hi + 1;
    }
    if (c0 < c1) {  // block-uniform
        const OffT p0 = rowPtr[c0], p1 = rowPtr[c1];
        constexpr uint32_t kIn = 8;  // loads in flight per thread
        for (OffT q0 = p0 + threadIdx.x; q0 < p1; q0 += kIn * kSplitBlock) {
            uint32_t id[kIn];
#pragma unroll
            for (uint32_t z = 0; z < kIn; z++) {
                const OffT q = q0 + (OffT)(z * kSplitBlock);
                id[z] = q < p1 ? colInd[q] : 0u;
            }
#pragma unroll
            for (uint32_t z = 0; z < kIn; z++)
                if (q0 + (OffT)(z * kSplitBlock) < p1) atomicAdd(&s_cnt[part_of(id[z])], 1u);
        }
        for (uint64_t v0 = c0 + threadIdx.x; v0 < c1; v0 += 4 * kSplitBlock) {
            uint32_t id[4];
            bool has[4];
#pragma unroll
            for (uint32_t z = 0; z < 4; z++) {
                const uint64_t v = v0 + z * kSplitBlock;
                OffT rb = 0, e = 0;
                if (v < c1) {
                    rb = rowPtr[v];
                    e = rowPtr[v + 1];
                }
                has[z] = e > rb;
                id[z] = has[z] ? colInd[e - 1] : 0u;
            }
#pragma unroll
            for (uint32_t z = 0; z < 4; z++)
                if (has[z]) atomicSub(&s_cnt[part_of(id[z])], 1u);
        }
        if (clip) {
            const uint64_t w_lo = S.w_lo, w_hi = S.w_hi;
            for (int z = 0; z < 2; z++) {
                const uint64_t ve = z == 0 ? v_lo : v_hi;
                if (ve < c0 || ve >= c1 || (z == 1 && v_hi == v_lo)) continue;  // block-uniform
                const OffT rb = rowPtr[ve], e = rowPtr[ve + 1];
                const uint64_t ws = S.ws_full[ve];
                for (OffT j = rb + threadIdx.x; j + 1 < e; j += kSplitBlock) {
                    OffT lo, hi;
                    clip_tail<OffT>(j, e, rb, ws, w_lo, w_hi, lo, hi);
                    if (lo >= hi) atomicSub(&s_cnt[part_of(colInd[j])], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < parts; i += kSplitBlock) S.hist1[b * parts + i] = s_cnt[i];
}

// Vertices with at least this many in-edges are taken by a whole wave.
constexpr uint32_t kSplitBigVertex = kWave;

// First level of the transposition.  Block b takes vertex chunk b (clipped
// to the counted vertices); a wave takes 64 consecutive vertices:
//  * a vertex with kSplitBigVertex or more in-edges is taken by the whole
//    wave in turn: the lanes load its ids 64 consecutive positions at a
//    time, four loads in flight;
//  * the others' in-edges are scanned and dealt to the lanes 64 at a time
//    (a lane finds its vertex by a binary lifting over the scanned counts),
//    four such rounds together so that their loads are in flight at once.
// Every in-edge (its tail clipped at the ends of the counted range; empty
// ones dropped) gets its key -- w - thr for a top row w, else rows +
// rowPtr[w] / kSlabSpan -- and belongs to the part key / kSplitPart.  With
// hist1 holding each part's records in earlier chunks (k_split_l1count,
// then the offset scans) and part_base its first record, each in-edge takes
// the next record slot of its part (an LDS cursor), and its key and record
// are stored there.  The count saw the same in-edges in the same chunks.
template <typename OffT>
__global__ __launch_bounds__(kSplitBlock) void k_split_l1(const GraphDesc *__restrict__ desc,
                                                          const SplitDesc *__restrict__ sd) {
    __shared__ uint32_t s_cnt[kSplitPartsMax];
    __shared__ uint64_t s_base[kSplitPartsMax];
    const GraphDesc &D = desc[0];
    const SplitDesc &S = sd[0];
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(D.rowPtr);
    const gptr<uint32_t> colInd = to_global(D.colInd);
    const uint64_t *__restrict__ wf = S.ws_full;
    uint64_t *__restrict__ hist1 = S.hist1;
    uint16_t *__restrict__ rec_id = S.rec_id;
    TailPair<OffT> *__restrict__ rec = static_cast<TailPair<OffT> *>(S.rec);
    uint32_t *__restrict__ rec_aux = S.rec_aux;
    const uint32_t n = D.n, thr = D.thr, rows = S.rows, parts = S.parts;
    const uint32_t b = blockIdx.x;
    for (uint32_t i = threadIdx.x; i < parts; i += kSplitBlock) {
        s_cnt[i] = 0u;
        s_base[i] = S.part_base[i] + hist1[b * parts + i];
    }
    __syncthreads();

    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x >> 6;

    // The in-edge of an act lane stored at its part's next record slot.  The lanes of a wave that share a part take consecutive
    // slots, and a chunk's few part cursors stay in the L2 cache, so the
    // stores leave in whole lines.
    auto take = [&](bool act, uint32_t key, OffT first, OffT hi, uint32_t aux) {
        if (!act) return;
        const uint32_t q = key / kSplitPart;
        const uint64_t p = s_base[q] + atomicAdd(&s_cnt[q], 1u);
        rec_id[p] = (uint16_t)(key & (kSplitPart - 1));
        rec[p] = TailPair<OffT>{first, hi};
        if (key >= rows) rec_aux[p] = aux;
    };
    // key of middle vertex id, given rowPtr[id] for a list row
    auto key_of = [&](uint32_t id, OffT rp) -> uint32_t {
        return id >= thr ? id - thr : rows + (uint32_t)((uint64_t)rp / kSlabSpan);
    };
    // slab_aux of a list row: adj+(id) = colInd[rp, rp1)
    auto aux_of = [&](OffT rp, OffT rp1) -> uint32_t {
        const uint64_t len = (uint64_t)(rp1 - rp);
        return (uint32_t)((uint64_t)rp % kSlabSpan) | ((len <= kSlabRowMax ? (uint32_t)len : kSlabAuxLong) << 13);
    };

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
        OffT e = 0, rb = 0;
        uint32_t cnt = 0;
        bool edge = false;
        uint64_t wsv = 0;
        if (v < c1) {
            rb = rowPtr[v];
            e = rowPtr[v + 1];
            cnt = (uint32_t)in_edges_of((uint64_t)(e - rb));
            edge = clip && (v == v_lo || v == v_hi);
            if (edge) wsv = wf[v];
        }
        // big vertices, one at a time by the whole wave
        uint64_t big = __ballot(cnt >= kSplitBigVertex);
        while (big) {
            const uint32_t src = (uint32_t)__builtin_ctzll(big);
            big &= big - 1;
            const OffT vrb = (OffT)readlane_u64((uint64_t)rb, src);
            const OffT vend = vrb + (OffT)(uint32_t)__builtin_amdgcn_readlane((int)cnt, (int)src);
            const OffT ve = (OffT)readlane_u64((uint64_t)e, src);
            const bool vedge = __builtin_amdgcn_readlane((int)edge, (int)src) != 0;
            const uint64_t vws = readlane_u64(wsv, src);
            for (OffT jb = vrb; jb < vend; jb += 4 * kWave) {
                uint32_t ids[4];
                OffT rp[4], rp1[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const OffT j = jb + (OffT)(u * kWave) + lane;
                    ids[u] = j < vend ? colInd[j] : 0u;
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const OffT j = jb + (OffT)(u * kWave) + lane;
                    rp[u] = rp1[u] = 0;
                    if (j < vend && ids[u] < thr) {
                        rp[u] = rowPtr[ids[u]];
                        rp1[u] = rowPtr[ids[u] + 1];
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const OffT j = jb + (OffT)(u * kWave) + lane;
                    OffT lo = j + 1, hi = ve;
                    if (vedge && j < vend) clip_tail<OffT>(j, ve, vrb, vws, w_lo, w_hi, lo, hi);
                    const bool top = ids[u] >= thr;
                    take(j < vend && lo < hi, key_of(ids[u], rp[u]), top ? lo : j, hi, aux_of(rp[u], rp1[u]));
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
        // kDeal rounds of 64 in-edges at a time, so that their colInd and
        // rowPtr loads are in flight together
        constexpr uint32_t kDeal = 4;
        for (uint32_t base = 0; base < tot; base += kDeal * kWave) {
            OffT j[kDeal], lo[kDeal], hi[kDeal];
            bool act[kDeal];
#pragma unroll
            for (uint32_t u = 0; u < kDeal; u++) {
                const uint32_t idx = base + u * kWave + lane;
                // the lane's vertex: the rightmost whose first in-edge is <= idx
                uint32_t pos = 0;
#pragma unroll
                for (uint32_t st = kWave / 2; st > 0; st >>= 1) {
                    const uint32_t c = pos + st;
                    const uint32_t x = (uint32_t)__shfl((int)excl, (int)(c & (kWave - 1)), kWave);
                    if (c < (uint32_t)kWave && x <= idx) pos = c;
                }
                const OffT rbv = shfl_off(rb, pos);
                j[u] = rbv + (OffT)(idx - (uint32_t)__shfl((int)excl, (int)pos, kWave));
                const OffT ev = shfl_off(e, pos);
                lo[u] = j[u] + 1;
                hi[u] = ev;
                if (any_edge) {
                    const bool ed = __shfl((int)edge, (int)pos, kWave) != 0;
                    const uint64_t wv0 = shfl_u64(wsv, pos);
                    if (ed) clip_tail<OffT>(j[u], ev, rbv, wv0, w_lo, w_hi, lo[u], hi[u]);
                }
                act[u] = idx < tot && lo[u] < hi[u];
            }
            uint32_t id[kDeal];
#pragma unroll
            for (uint32_t u = 0; u < kDeal; u++) id[u] = act[u] ? colInd[j[u]] : 0u;
            OffT rp[kDeal], rp1[kDeal];
#pragma unroll
            for (uint32_t u = 0; u < kDeal; u++) {
                rp[u] = rp1[u] = 0;
                if (act[u] && id[u] < thr) {
                    rp[u] = rowPtr[id[u]];
                    rp1[u] = rowPtr[id[u] + 1];
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < kDeal; u++)
                take(act[u], key_of(id[u], rp[u]), id[u] >= thr ? lo[u] : j[u], hi[u], aux_of(rp[u], rp1[u]));
        }
    }
}

// Exclusive scan over the kSplitBlock threads of a block; the block's sum
// into total.  s_w holds a word per wave.
__device__ __forceinline__ uint64_t split_block_scan(uint64_t x, uint64_t *s_w, uint64_t &total) {
    constexpr int kWaves = kSplitBlock / kWave;
    const uint64_t inc = wave_incl_scan(x);
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x >> 6;
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
    total = all;
    return off + inc - x;
}

// Block q (a part), thread c (a chunk): hist1[c][q] becomes the part's
// records in earlier chunks, and part_tot[q] its total.
__global__ __launch_bounds__(kSplitBlock) void k_split_l1cols(const SplitDesc *__restrict__ sd) {
    __shared__ uint64_t s_w[kSplitBlock / kWave];
    const SplitDesc &S = sd[0];
    const uint32_t q = blockIdx.x, c = threadIdx.x, parts = S.parts;
    uint64_t *__restrict__ hist1 = S.hist1;
    const uint64_t x = c < S.chunks ? hist1[(uint64_t)c * parts + q] : 0;
    uint64_t tot;
    const uint64_t ex = split_block_scan(x, s_w, tot);
    if (c < S.chunks) hist1[(uint64_t)c * parts + q] = ex;
    if (c == 0) S.part_tot[q] = tot;
}

// One block, thread per part: part_base and c2base (kSplitChunk2 records per
// second-level block) as prefixes of the part totals.
__global__ __launch_bounds__(kBlock) void k_split_l1bases(const SplitDesc *__restrict__ sd) {
    __shared__ uint64_t s_tot[kWavesPerBlock];
    static_assert(kSplitPartsMax <= (uint32_t)kBlock, "a part per thread");
    const SplitDesc &S = sd[0];
    const uint32_t q = threadIdx.x, parts = S.parts;
    const uint64_t t = q < parts ? S.part_tot[q] : 0, nb = (t + kSplitChunk2 - 1) / kSplitChunk2;
    uint64_t tt, tb;
    const uint64_t et = block_excl_scan(t, s_tot, tt);
    const uint64_t eb = block_excl_scan(nb, s_tot, tb);
    if (q < parts) {
        S.part_base[q] = et;
        S.c2base[q] = (uint32_t)eb;
    }
    if (q == 0) {
        S.part_base[parts] = tt;
        S.c2base[parts] = (uint32_t)tb;
    }
}

// Second and third levels.  Level 2 groups each part's A records by
// sub-part into B; level 3 groups each sub-part's B records by key into ie.
// A level's input bucket b (a part, or a sub-part) holds records
// [base_in[b], base_in[b+1]) and is split into blocks of kChunk records
// (its blocks start at cbase[b]); an output bucket is one of the kF
// sub-parts of a part, or one of the kSplitSub keys of a sub-part.
// k_split_level_count counts a block's records per output bucket into
// hist[c][0..kF); k_split_level_scatter (hist then holding, per output
// bucket, the records of the input bucket's earlier blocks) places each
// record at the next slot of its output bucket.  Blocks stride over the
// level's blocks.
constexpr int kLevelRecords = 4;  // records per thread per step, loads issued together
// Most blocks a level launches (they stride over its blocks of records).
constexpr uint32_t kLevelGridMax = 2048;

// A level's input keys are the keys within their input bucket (15 bits in
// A, 8 in B): key = b * kUnit + local.
template <int kLevel>
struct SplitLevel {
    static constexpr uint32_t kF = kLevel == 2 ? kSplitFan2 : kSplitSub;
    static constexpr uint32_t kShift = kLevel == 2 ? kSplitSubShift : 0u;
    static constexpr uint32_t kChunk = kLevel == 2 ? kSplitChunk2 : kSplitChunk3;
    static constexpr uint32_t kUnit = kLevel == 2 ? kSplitPart : kSplitSub;
    using KeyIn = typename std::conditional<kLevel == 2, uint16_t, uint8_t>::type;
};

// The last b in [0, nb) with cbase[b] <= c (cbase[0] = 0 <= c).
__device__ __forceinline__ uint32_t bucket_of_block(const uint32_t *cbase, uint32_t nb, uint32_t c) {
    uint32_t lo = 0, hi = nb;  // cbase[lo] <= c < cbase[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cbase[mid] <= c) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <typename OffT, int kLevel>
__global__ __launch_bounds__(kSplitBlock) void k_split_level_count(const SplitDesc *__restrict__ sd) {
    using L = SplitLevel<kLevel>;
    constexpr uint32_t kF = L::kF;
    __shared__ uint32_t s_cnt[kF];
    const SplitDesc &S = sd[0];
    const uint32_t nb = kLevel == 2 ? S.parts : S.subs;
    const uint64_t *__restrict__ base_in = kLevel == 2 ? S.part_base : S.sub_base;
    const uint32_t *__restrict__ cbase = kLevel == 2 ? S.c2base : S.c3base;
    uint32_t *__restrict__ hist = kLevel == 2 ? S.hist2 : S.hist3;
    const typename L::KeyIn *__restrict__ key_in =
        reinterpret_cast<const typename L::KeyIn *>(kLevel == 2 ? (const void *)S.rec_id : (const void *)S.rec_id2);
    const uint32_t nblk = cbase[nb];
    for (uint32_t c = blockIdx.x; c < nblk; c += gridDim.x) {
        const uint32_t b = bucket_of_block(cbase, nb, c);
        const uint64_t k0 = base_in[b] + (uint64_t)(c - cbase[b]) * L::kChunk;
        const uint64_t kq = base_in[b + 1], k1 = k0 + L::kChunk < kq ? k0 + L::kChunk : kq;
        __syncthreads();  // the previous block's LDS reads are done
        for (uint32_t i = threadIdx.x; i < kF; i += kSplitBlock) s_cnt[i] = 0u;
        __syncthreads();
        constexpr int kU = kLevelRecords;
        for (uint64_t kb = k0 + threadIdx.x; kb < k1; kb += (uint64_t)kU * kSplitBlock) {
            uint32_t key[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const uint64_t k = kb + (uint64_t)u * kSplitBlock;
                key[u] = k < k1 ? key_in[k] : kSentinel;
            }
#pragma unroll
            for (int u = 0; u < kU; u++)
                if (key[u] != kSentinel) atomicAdd(&s_cnt[(key[u] >> L::kShift) & (kF - 1)], 1u);
        }
        __syncthreads();
        uint32_t *__restrict__ h = hist + (uint64_t)c * kF;
        for (uint32_t i = threadIdx.x; i < kF; i += kSplitBlock) h[i] = s_cnt[i];
    }
}

// Records a level's scatter stages in LDS per tile: 4096 with 4-byte
// offsets, 2048 with 8-byte ones (32 + 8 KB of pairs and words either way).
template <typename OffT>
struct LevelTile {
    static constexpr uint32_t kT = sizeof(OffT) == 4 ? 4096 : 2048;
};

// The scatter of a level.  hist[c] holds, per output bucket, the input
// bucket's records in blocks before c.  The block takes its records in tiles
// of kT: each tile is ranked per output bucket (LDS atomics), staged in LDS
// sorted by output bucket, and written out in that order, so that the
// records of one output bucket in a tile -- kT / kF of them on average --
// leave as consecutive lanes of one store to consecutive addresses.
template <typename OffT, int kLevel>
__global__ __launch_bounds__(kSplitBlock) void k_split_level_scatter(const SplitDesc *__restrict__ sd) {
    using L = SplitLevel<kLevel>;
    constexpr uint32_t kF = L::kF;
    constexpr uint32_t kT = LevelTile<OffT>::kT;
    constexpr uint32_t kPerThread = kT / kSplitBlock;
    __shared__ uint64_t s_base[kF];  // next output slot of each output bucket
    __shared__ uint32_t s_tcnt[kF];  // the tile's records per output bucket
    __shared__ uint32_t s_toff[kF];  // their exclusive prefix: each bucket's first staged slot
    __shared__ uint32_t st_key[kT];
    __shared__ TailPair<OffT> st_rec[kT];
    __shared__ uint32_t st_aux[kT];
    const SplitDesc &S = sd[0];
    const uint32_t rows = S.rows, keys = S.keys;
    const uint32_t nb = kLevel == 2 ? S.parts : S.subs;
    const uint64_t *__restrict__ base_in = kLevel == 2 ? S.part_base : S.sub_base;
    const uint32_t *__restrict__ cbase = kLevel == 2 ? S.c2base : S.c3base;
    const uint32_t *__restrict__ hist = kLevel == 2 ? S.hist2 : S.hist3;
    const typename L::KeyIn *__restrict__ key_in =
        reinterpret_cast<const typename L::KeyIn *>(kLevel == 2 ? (const void *)S.rec_id : (const void *)S.rec_id2);
    const TailPair<OffT> *__restrict__ rec_in = static_cast<const TailPair<OffT> *>(kLevel == 2 ? S.rec : S.rec2);
    const uint32_t *__restrict__ aux_in = kLevel == 2 ? S.rec_aux : S.rec_aux2;
    uint8_t *__restrict__ key_out = S.rec_id2;  // level 2 only
    TailPair<OffT> *__restrict__ rec_out = static_cast<TailPair<OffT> *>(kLevel == 2 ? S.rec2 : S.ie);
    uint32_t *__restrict__ aux_out = kLevel == 2 ? S.rec_aux2 : S.ie_aux;
    const OffT *__restrict__ rp_in = static_cast<const OffT *>(S.rp_in);
    const uint32_t lane = threadIdx.x & (kWave - 1);
    auto bucket = [](uint32_t key) { return (key >> L::kShift) & (kF - 1); };
    const uint32_t nblk = cbase[nb];
    for (uint32_t c = blockIdx.x; c < nblk; c += gridDim.x) {
        const uint32_t b = bucket_of_block(cbase, nb, c);
        const uint64_t k0 = base_in[b] + (uint64_t)(c - cbase[b]) * L::kChunk;
        const uint64_t kq = base_in[b + 1], k1 = k0 + L::kChunk < kq ? k0 + L::kChunk : kq;
        const uint32_t *__restrict__ h = hist + (uint64_t)c * kF;
        __syncthreads();  // the previous block's LDS reads are done
        for (uint32_t i = threadIdx.x; i < kF; i += kSplitBlock) {
            // first slot of output bucket i: a sub-part's first record, or a key's first in-edge
            const uint32_t o = b * kF + i;
            const uint64_t ob = kLevel == 2 ? S.sub_base[o] : o < keys ? (uint64_t)rp_in[o] : 0;
            s_base[i] = ob + h[i];
        }
        for (uint64_t t0 = k0; t0 < k1; t0 += kT) {  // block-uniform
            for (uint32_t i = threadIdx.x; i < kF; i += kSplitBlock) s_tcnt[i] = 0u;
            __syncthreads();
            uint32_t key[kPerThread], rank[kPerThread], aux[kPerThread];
            TailPair<OffT> t[kPerThread];
#pragma unroll
            for (uint32_t u = 0; u < kPerThread; u++) {
                const uint64_t k = t0 + threadIdx.x + (uint64_t)u * kSplitBlock;
                key[u] = k < k1 ? key_in[k] : kSentinel;
                aux[u] = 0u;
                if (key[u] != kSentinel) {
                    t[u] = rec_in[k];
                    if (b * L::kUnit + key[u] >= rows) aux[u] = aux_in[k];  // a list record
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < kPerThread; u++)
                if (key[u] != kSentinel) rank[u] = atomicAdd(&s_tcnt[bucket(key[u])], 1u);
            __syncthreads();
            if (threadIdx.x < kWave) {  // exclusive prefix of the tile's bucket counts, kF / 64 per lane
                constexpr uint32_t kPer = kF / kWave;
                uint32_t v[kPer], sum = 0;
#pragma unroll
                for (uint32_t z = 0; z < kPer; z++) {
                    v[z] = s_tcnt[lane * kPer + z];
                    sum += v[z];
                }
                uint32_t incl = sum;
#pragma unroll
                for (int o = 1; o < kWave; o <<= 1) {
                    const uint32_t y = (uint32_t)__shfl_up((int)incl, o, kWave);
                    if ((int)lane >= o) incl += y;
                }
                uint32_t run = incl - sum;
#pragma unroll
                for (uint32_t z = 0; z < kPer; z++) {
                    s_toff[lane * kPer + z] = run;
                    run += v[z];
                }
            }
            __syncthreads();
#pragma unroll
            for (uint32_t u = 0; u < kPerThread; u++) {
                if (key[u] == kSentinel) continue;
                const uint32_t at = s_toff[bucket(key[u])] + rank[u];
                st_key[at] = key[u];
                st_rec[at] = t[u];
                st_aux[at] = aux[u];
            }
            __syncthreads();
            const uint32_t nt = (uint32_t)(k1 - t0 < kT ? k1 - t0 : kT);
#pragma unroll
            for (uint32_t u = 0; u < kPerThread; u++) {
                const uint32_t i = threadIdx.x + u * kSplitBlock;
                if (i >= nt) continue;
                const uint32_t k = st_key[i], q = bucket(k);  // the key within the input bucket
                const uint64_t p = s_base[q] + (i - s_toff[q]);
                if (kLevel == 2) key_out[p] = (uint8_t)(k & (kSplitSub - 1));
                rec_out[p] = st_rec[i];
                if (b * L::kUnit + k >= rows) aux_out[p] = st_aux[i];
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < kF; i += kSplitBlock) s_base[i] += s_tcnt[i];
        }
    }
}

// Thread per output bucket o = b*kF + i of a level: the column prefix of
// hist over input bucket b's blocks (hist[c][i] becomes o's records in b's
// earlier blocks) and o's total (sub_tot, or key_tot at level 3).
template <int kLevel>
__global__ __launch_bounds__(kBlock) void k_split_levelcols(const SplitDesc *__restrict__ sd) {
    constexpr uint32_t kF = SplitLevel<kLevel>::kF;
    const SplitDesc &S = sd[0];
    const uint32_t o = blockIdx.x * kBlock + threadIdx.x;
    if (o >= (kLevel == 2 ? S.subs : S.keys)) return;
    const uint32_t b = o / kF, i = o % kF;
    const uint32_t *__restrict__ cbase = kLevel == 2 ? S.c2base : S.c3base;
    uint32_t *__restrict__ hist = kLevel == 2 ? S.hist2 : S.hist3;
    uint32_t s = 0;
    const uint32_t c1 = cbase[b + 1];
    for (uint32_t c = cbase[b]; c < c1; c += 8) {  // eight loads in flight
        uint32_t x[8];
#pragma unroll
        for (uint32_t z = 0; z < 8; z++) x[z] = c + z < c1 ? hist[(uint64_t)(c + z) * kF + i] : 0u;
#pragma unroll
        for (uint32_t z = 0; z < 8; z++) {
            if (c + z < c1) hist[(uint64_t)(c + z) * kF + i] = s;
            s += x[z];
        }
    }
    (kLevel == 2 ? S.sub_tot : S.key_tot)[o] = s;
}

// The second level's column prefixes, as k_split_levelcols<2>, with one
// block per part: the part's blocks are cut into kSplitBlock / kSplitFan2
// segments whose column sums are taken in parallel first (a part spans
// hundreds of blocks, which one thread per column walked in turn).
__global__ __launch_bounds__(kSplitBlock) void k_split_levelcols2(const SplitDesc *__restrict__ sd) {
    constexpr uint32_t kF = kSplitFan2, kSeg = kSplitBlock / kF;
    __shared__ uint32_t s_sum[kSeg][kF];
    const SplitDesc &S = sd[0];
    const uint32_t b = blockIdx.x, i = threadIdx.x % kF, sg = threadIdx.x / kF;
    const uint32_t cb0 = S.c2base[b], nb = S.c2base[b + 1] - cb0;
    const uint32_t per = (nb + kSeg - 1) / kSeg;
    const uint32_t s0 = cb0 + (sg * per < nb ? sg * per : nb), s1 = cb0 + ((sg + 1) * per < nb ? (sg + 1) * per : nb);
    uint32_t *__restrict__ hist = S.hist2;
    uint32_t sum = 0;
#pragma unroll 8
    for (uint32_t c = s0; c < s1; c++) sum += hist[(uint64_t)c * kF + i];
    s_sum[sg][i] = sum;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (uint32_t z = 0; z < kSeg; z++) {
        const uint32_t x = s_sum[z][i];
        pre += z < sg ? x : 0u;
        tot += x;
    }
#pragma unroll 8
    for (uint32_t c = s0; c < s1; c++) {
        const uint32_t x = hist[(uint64_t)c * kF + i];
        hist[(uint64_t)c * kF + i] = pre;
        pre += x;
    }
    if (sg == 0) S.sub_tot[b * kF + i] = tot;
}

// One block: sub_base = exclusive prefix of the sub-part totals (the
// sub-parts in key order, so each part's sub-parts start at its own base)
// and c3base = exclusive prefix of their third-level blocks.
__global__ __launch_bounds__(kSplitBlock) void k_split_subscan(const SplitDesc *__restrict__ sd) {
    constexpr int kWaves = kSplitBlock / kWave;
    __shared__ uint64_t s_r[kWaves], s_b[kWaves];
    const SplitDesc &S = sd[0];
    const uint32_t subs = S.subs;
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x >> 6;
    uint64_t cr = 0, cb = 0;
    for (uint32_t base = 0; base < subs; base += kSplitBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint64_t x = i < subs ? S.sub_tot[i] : 0;
        const uint64_t y = (x + kSplitChunk3 - 1) / kSplitChunk3;
        const uint64_t ix = wave_incl_scan(x), iy = wave_incl_scan(y);
        if (lane == kWave - 1) {
            s_r[wv] = ix;
            s_b[wv] = iy;
        }
        __syncthreads();
        uint64_t ox = 0, oy = 0, ax = 0, ay = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) {
            ox += w < (int)wv ? s_r[w] : 0;
            oy += w < (int)wv ? s_b[w] : 0;
            ax += s_r[w];
            ay += s_b[w];
        }
        __syncthreads();
        if (i < subs) {
            S.sub_base[i] = cr + ox + ix - x;
            S.c3base[i] = (uint32_t)(cb + oy + iy - y);
        }
        cr += ax;
        cb += ay;
    }
    if (threadIdx.x == 0) {
        S.sub_base[subs] = cr;
        S.c3base[subs] = (uint32_t)cb;
    }
}

// Block b: the in-edges of keys [b*kKeyChunk, (b+1)*kKeyChunk) into
// key_part[b].
__global__ __launch_bounds__(kSplitBlock) void k_split_keysum(const SplitDesc *__restrict__ sd) {
    __shared__ uint64_t s_w[kSplitBlock / kWave];
    const SplitDesc &S = sd[0];
    const uint64_t i0 = (uint64_t)blockIdx.x * kKeyChunk + 4ull * threadIdx.x;
    uint64_t x = 0;
#pragma unroll
    for (uint32_t z = 0; z < 4; z++)
        if (i0 + z < S.keys) x += S.key_tot[i0 + z];
    x = wave_sum(x);
    if ((threadIdx.x & (kWave - 1)) == 0) s_w[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < kSplitBlock / kWave; w++) t += s_w[w];
        S.key_part[blockIdx.x] = t;
    }
}

// Block b (key_part then holding the exclusive prefix of the key blocks):
// rp_in of its keys; the block of key `rows` puts the top rows' in-edges
// into meta, the last block rp_in[keys] and the rest of meta (tile counters
// and wedge tallies zeroed, the counted wedges and all in-edges) and
// counts[0] = 0.
template <typename OffT>
__global__ __launch_bounds__(kSplitBlock) void k_split_keyscan(const SplitDesc *__restrict__ sd,
                                                               uint64_t *__restrict__ meta,
                                                               unsigned long long *__restrict__ counts) {
    __shared__ uint64_t s_w[kSplitBlock / kWave];
    const SplitDesc &S = sd[0];
    const uint32_t keys = S.keys, rows = S.rows;
    OffT *__restrict__ rp = static_cast<OffT *>(S.rp_in);
    const uint64_t i0 = (uint64_t)blockIdx.x * kKeyChunk + 4ull * threadIdx.x;
    uint32_t v[4];
    uint64_t x = 0;
#pragma unroll
    for (uint32_t z = 0; z < 4; z++) {
        v[z] = i0 + z < keys ? S.key_tot[i0 + z] : 0u;
        x += v[z];
    }
    uint64_t tot;
    uint64_t at = S.key_part[blockIdx.x] + split_block_scan(x, s_w, tot);
#pragma unroll
    for (uint32_t z = 0; z < 4; z++) {
        if (i0 + z < keys) rp[i0 + z] = (OffT)at;
        if (i0 + z == rows) meta[kMetaInEdges] = at;
        at += v[z];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        const uint64_t all = S.key_part[blockIdx.x] + tot;
        rp[keys] = (OffT)all;
        if (rows >= keys) meta[kMetaInEdges] = all;
        meta[kMetaRows] = 0;
        meta[kMetaSlabs] = 0;
        meta[kMetaTopWedges] = 0;
        meta[kMetaListWedges] = 0;
        meta[kMetaWedges] = S.w_hi - S.w_lo;
        meta[kMetaAllEdges] = all;
        counts[0] = 0;
    }
}

// ------------------------------------------------------------- top-a part

// The top rows' tails as packed ids.  A top in-edge v -> w (w >= thr) has a
// tail of ids above w, so each of them lies in [thr, n) and its offset x =
// u - thr is below rows (the a-bit row must fit a block's LDS, so rows <=
// 2^21).  The packed array is a sequence of 16-byte chunks; chunk c holds
// the fields of colInd positions [Pc, Pc + P) in bits [Bi, Bi + B), i < P,
// with the fewest bits B the graph's rows allow: the row kernel streams 16/P
// bytes per tail id instead of 4.  Layouts (PackLayout<L>):
//   0: P = 6, B = 21, rows <= 2^19, field (x >> 5) << 7 | (x & 31): bits
//      [5, 21) are the LDS byte address of x's bitmap word and bits [0, 5)
//      its bit, so that the row kernel takes both with one extract;
//   1: P = 6, B = 21, rows <= 2^21, field x;
//   2: P = 7, B = 18, rows <= 2^18, field x;
//   3: P = 8, B = 16, rows <= 2^16, field x.
// An id below thr (never in a top tail) packs as 0.
template <int L>
struct PackLayout;
template <>
struct PackLayout<0> {
    static constexpr int P = 6, B = 21;
    static constexpr bool kAddr = true;
};
template <>
struct PackLayout<1> {
    static constexpr int P = 6, B = 21;
    static constexpr bool kAddr = false;
};
template <>
struct PackLayout<2> {
    static constexpr int P = 7, B = 18;
    static constexpr bool kAddr = false;
};
template <>
struct PackLayout<3> {
    static constexpr int P = 8, B = 16;
    static constexpr bool kAddr = false;
};
constexpr uint32_t kPackRowsMax = 1u << 21;
constexpr uint32_t kPackGrid = 4096;  // blocks of the packing kernel (grid-stride)

// The layout of a graph with `rows` top rows (rows <= kPackRowsMax).
__host__ __device__ __forceinline__ int pack_layout(uint32_t rows) {
    return rows <= (1u << 16) ? 3 : rows <= (1u << 18) ? 2 : rows <= (1u << 19) ? 0 : 1;
}

// Field i of a chunk, shifted down so that its bits start at bit 0 (bits
// above B hold the next field's; the readers extract what they need).
template <int B, int i>
__device__ __forceinline__ uint32_t chunk_field(u32x4 d) {
    constexpr int bit = B * i, k = bit / 32, sh = bit % 32;
    if constexpr (sh + B <= 32) {
        return d[k] >> sh;
    } else {
        return __builtin_amdgcn_alignbit(d[k + 1], d[k], sh);
    }
}

template <int B, int... I>
__device__ __forceinline__ void chunk_fields(u32x4 d, uint32_t *f, std::integer_sequence<int, I...>) {
    ((f[I] = chunk_field<B, I>(d)), ...);
}

// Packs colInd into S.rec2 (free once the third level has grouped its
// records): a block takes kBlock chunks at a time, loads their P * kBlock
// ids into LDS with coalesced 4-byte loads, and each thread packs one chunk
// from LDS (stride P words: no bank conflicts for odd P, two-way for P = 6
// and 8) into one 16-byte store.  Positions past m pack as 0.  Writes
// ceil(m / P) * 16 bytes, within rec2's cap * 2o bytes (and the 256-byte
// rounding of every workspace piece).
template <int L>
__device__ __forceinline__ void pack_chunks(const GraphDesc &D, uint64_t m, u32x4 *__restrict__ dst,
                                            uint32_t *s_ids) {
    using PL = PackLayout<L>;
    constexpr int P = PL::P, B = PL::B;
    const gptr<uint32_t> colInd = to_global(D.colInd);
    const uint32_t thr = D.thr;
    const uint64_t chunks = (m + P - 1) / P;
    for (uint64_t g0 = (uint64_t)blockIdx.x * kBlock; g0 < chunks; g0 += (uint64_t)gridDim.x * kBlock) {
        const uint64_t q0 = g0 * P;
#pragma unroll
        for (int z = 0; z < P; z++) {
            const uint64_t q = q0 + (uint64_t)z * kBlock + threadIdx.x;
            s_ids[z * kBlock + threadIdx.x] = q < m ? colInd[q] : 0u;
        }
        __syncthreads();
        const uint64_t t = g0 + threadIdx.x;
        if (t < chunks) {
            uint32_t d[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int i = 0; i < P; i++) {
                const uint32_t u = s_ids[P * threadIdx.x + i];
                const uint32_t x = u >= thr ? u - thr : 0u;
                const uint32_t f = PL::kAddr ? ((x >> 5) << 7) | (x & 31u) : x;
                const int bit = B * i, k = bit / 32, sh = bit % 32;
                d[k] |= f << sh;
                if (sh + B > 32) d[k + 1] |= f >> (32 - sh);
            }
            dst[t] = u32x4{d[0], d[1], d[2], d[3]};
        }
        __syncthreads();
    }
}

template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_split_pack(const GraphDesc *__restrict__ desc,
                                                       const SplitDesc *__restrict__ sd) {
    __shared__ uint32_t s_ids[8 * kBlock];
    const GraphDesc &D = desc[0];
    const SplitDesc &S = sd[0];
    const uint64_t m = (uint64_t) static_cast<const OffT *>(D.rowPtr)[D.n];
    u32x4 *__restrict__ dst = static_cast<u32x4 *>(S.rec2);
    switch (pack_layout(S.rows)) {  // block-uniform
        case 0: pack_chunks<0>(D, m, dst, s_ids); break;
        case 1: pack_chunks<1>(D, m, dst, s_ids); break;
        case 2: pack_chunks<2>(D, m, dst, s_ids); break;
        default: pack_chunks<3>(D, m, dst, s_ids); break;
    }
}

// Lanes per in-edge tail in the row kernel (a wave takes kWave / kSplitLanes
// in-edges per step) and 16-byte chunks per lane per step.  Tails average
// ~180 ids on R-MAT-24 (median 86), and a step lasts as long as its longest
// tail: 8 lanes of one chunk keep 80% of the id slots busy, 16 lanes of two
// chunks 57%.  Measured on R-MAT-24: 18.2 ms at 8 x 1, 19.0 at 8 x 2, 18.8
// at 16 x 1, 19.1 at 4 x 1, 22.1 at 16 x 2.
constexpr int kSplitLanes = 8;
constexpr int kSplitChunks = 1;

// Counts the hits of the P packed ids of chunk c of a tail.  base = Pc - lo
// (mod 2^32): position Pc + z is in the tail iff base + z < len.  Every
// packed field names a bit of the row bitmap, so all P are looked up (word
// and bit straight from the fields: a field's bit offset is its low 5
// bits), the hits gathered as a P-bit mask and the positions outside the
// tail masked off at the end.  An address field's bits [5, 21) are the byte
// offset of its word in the bitmap.
template <int L>
__device__ __forceinline__ uint32_t rows_chunk_hits(const uint32_t *s_bits, u32x4 ch, uint32_t base,
                                                    uint32_t len) {
    using PL = PackLayout<L>;
    constexpr int P = PL::P, B = PL::B;
    uint32_t f[P];
    chunk_fields<B>(ch, f, std::make_integer_sequence<int, P>{});
    uint32_t wrd[P];
#pragma unroll
    for (int z = 0; z < P; z++) {
        if (PL::kAddr) {
            const uint32_t addr = __builtin_amdgcn_ubfe(f[z], 5u, 16u);
            wrd[z] = *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(s_bits) + addr);
        } else {
            wrd[z] = s_bits[__builtin_amdgcn_ubfe(f[z], 5u, (uint32_t)B - 5u)];
        }
    }
    uint32_t hm = 0;
#pragma unroll
    for (int z = 0; z < P; z++) hm |= __builtin_amdgcn_ubfe(wrd[z], f[z], 1u) << z;
    // valid z: [lc, hc) with lc = -base when base is negative (the tail's
    // first chunk), hc = min(len - base, P)
    const uint32_t lc = base > 0xfffffff0u ? 0u - base : 0u;
    const uint32_t rem = len - base, hc = rem < (uint32_t)P ? rem : (uint32_t)P;
    const uint32_t vm = ((1u << hc) - 1u) & (0xffffffffu << lc);
    return __builtin_popcount(hm & vm);
}

// Persistent blocks take tiles of kSplitTile consecutive in-edges (in row
// order) from one counter.  Per row of a tile: adj+(w) set as bits (u - thr)
// in LDS, the tile's in-edges of the row dealt to the waves (kLanes lanes
// per in-edge), and each tail streamed through the bitmap from the packed
// ids (k_split_pack) by 16-byte chunks of P ids (6 to 8, PackLayout), kCh
// chunks per lane per step and the next step's chunks loaded before this
// step's are counted; then the row's bits cleared.
template <typename OffT, int kWaves, int kLanes, int kCh, int kLayout>
__global__ __launch_bounds__(kWave * kWaves) void k_split_rows(const GraphDesc *__restrict__ desc,
                                                               const SplitDesc *__restrict__ sd,
                                                               uint64_t *__restrict__ meta,
                                                               unsigned long long *__restrict__ counts) {
    // dynamic LDS only: the row bitmap, then the tile slot
    extern __shared__ uint32_t s_bits[];
    const GraphDesc &D = desc[0];
    const SplitDesc &S = sd[0];
    const cptr<OffT> rowPtrU = to_const(static_cast<const OffT *>(D.rowPtr));
    const gptr<uint32_t> colInd = to_global(D.colInd);
    const gptr<u32x4> packed = to_global(static_cast<const u32x4 *>(S.rec2));
    const cptr<OffT> rp_in = to_const(static_cast<const OffT *>(S.rp_in));
    const cptr<OffT> ieU = to_const(static_cast<const OffT *>(S.ie));
    const gptr<OffT> ieG = to_global(static_cast<const OffT *>(S.ie));
    const uint32_t thr = D.thr, rows = S.rows;
    constexpr uint32_t kPer = kWave / kLanes;  // in-edges a wave takes per step
    const uint32_t words = (rows + 31) >> 5;
    uint64_t &s_tile = *reinterpret_cast<uint64_t *>(s_bits + ((words + 1) & ~1u));
    for (uint32_t i = threadIdx.x; i < words; i += kWave * kWaves) s_bits[i] = 0;
    const uint64_t K = (uint64_t)rp_in[rows];
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = uni32(threadIdx.x >> 6);
    const uint32_t sub = lane / kLanes, sl = lane % kLanes;
    unsigned long long *const ctr = reinterpret_cast<unsigned long long *>(meta + kMetaRows);
    uint64_t hits = 0, checked = 0;
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
                if (sl == 0) checked += (uint64_t)(hi - lo);
                // chunk c holds positions [Pc, Pc + P); a tail is shorter
                // than 2^32 positions
                constexpr uint32_t P = PackLayout<kLayout>::P;
                const uint32_t len = (uint32_t)(hi - lo);
                const OffT c_end = hi / P + (hi % P != 0 ? 1 : 0);
                OffT c = lo / P + sl;
                u32x4 cur[kCh];
#pragma unroll
                for (int q = 0; q < kCh; q++) {
                    cur[q] = u32x4{0u, 0u, 0u, 0u};
                    if (c + (OffT)(q * kLanes) < c_end) cur[q] = packed[c + (OffT)(q * kLanes)];
                }
                while (c < c_end) {
                    const OffT cn = c + (OffT)(kCh * kLanes);
                    // the next chunks, or chunk c again past the tail's end
                    // (loaded unconditionally; a chunk past c_end is not counted)
                    u32x4 nxt[kCh];
#pragma unroll
                    for (int q = 0; q < kCh; q++) {
                        const OffT cq = cn + (OffT)(q * kLanes);
                        nxt[q] = packed[cq < c_end ? cq : c];
                    }
                    const uint32_t base = (uint32_t)(P * c - lo);
#pragma unroll
                    for (int q = 0; q < kCh; q++)
                        if (q == 0 || c + (OffT)(q * kLanes) < c_end)
                            h += rows_chunk_hits<kLayout>(s_bits, cur[q], base + P * q * kLanes, len);
#pragma unroll
                    for (int q = 0; q < kCh; q++) cur[q] = nxt[q];
                    c = cn;
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
    checked = wave_sum(checked);
    if (lane == 0 && hits) atomicAdd(&counts[0], (unsigned long long)hits);
    if (lane == 0 && checked) atomicAdd(reinterpret_cast<unsigned long long *>(meta + kMetaTopWedges), checked);
}

// ---------------------------------------------------------------- list part

// Bucket of id u in a row's region of len buckets (multiplicative hash).
__device__ __forceinline__ uint32_t slab_bucket(uint32_t u, uint32_t len) {
    return __umulhi(u * 0x9e3779b1u, len);
}

// Persistent blocks take tiles of kSlabTile consecutive list in-edges (in
// slab order) from one counter.  A tile's in-edges of one slab form a
// segment.  Per segment, the block hashes into LDS every row of the slab the
// segment's in-edges name that is not there yet (the table is kept while
// the next segment stays in the slab; a wave per row clears its buckets and
// inserts its ids): the row starting at window position st with len
// entries owns the buckets [st, st + len) of kSlabWays slots each, so a
// quarter of its slots are filled and a lookup nearly always ends in its
// first bucket (one 16-byte LDS read).  Then the segment's in-edges are
// dealt to the waves, kLanes lanes each; each tail is streamed by 16-byte
// loads, two in flight per lane, and each id u looked up in adj+(w) by
// linear probing from its hash bucket, or by a binary search in HBM for a
// row longer than kSlabRowMax.
// Same alignment and slack needs as the row kernel.
template <typename OffT, int kWaves, int kLanes>
__global__ __launch_bounds__(kWave * kWaves) void k_split_slabs(const GraphDesc *__restrict__ desc,
                                                                const SplitDesc *__restrict__ sd,
                                                                uint64_t *__restrict__ meta,
                                                                unsigned long long *__restrict__ counts) {
    constexpr uint32_t kRefWords = kSlabSpan / 32;
    __shared__ __align__(16) uint32_t s_tab[kSlabWin * kSlabWays];
    __shared__ uint32_t s_ref[kRefWords];  // rows (by start position) hashed so far
    __shared__ uint32_t s_new[kSlabTile];  // slab_aux words of the rows a segment adds
    __shared__ uint32_t s_nnew;
    __shared__ uint64_t s_tile;
    const GraphDesc &D = desc[0];
    const SplitDesc &S = sd[0];
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(D.rowPtr);
    const gptr<uint32_t> colInd = to_global(D.colInd);
    const cptr<OffT> rp_in = to_const(static_cast<const OffT *>(S.rp_in));
    const gptr<OffT> ieG = to_global(static_cast<const OffT *>(S.ie));
    const gptr<uint32_t> auxG = to_global(static_cast<const uint32_t *>(S.ie_aux));
    const uint32_t rows = S.rows, keys = S.keys;
    const uint64_t L0 = (uint64_t)rp_in[rows], L1 = (uint64_t)rp_in[keys];
    const OffT clip_j = (OffT)S.clip_j, clip_lo = (OffT)S.clip_lo;
    const bool clipped = S.clip_j != ~0ull;
    constexpr uint32_t kThreads = kWave * kWaves;
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = uni32(threadIdx.x >> 6);
    unsigned long long *const ctr = reinterpret_cast<unsigned long long *>(meta + kMetaSlabs);
    uint32_t cur = kSentinel;  // slab in s_tab (block-uniform)
    uint64_t hits = 0, checked = 0;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_tile = atomicAdd(ctr, 1ull);
        __syncthreads();
        const uint64_t k0 = L0 + uni64(s_tile) * kSlabTile;
        if (k0 >= L1) break;  // block-uniform
        const uint64_t k1 = k0 + kSlabTile < L1 ? k0 + kSlabTile : L1;
        uint64_t k = k0;
        while (k < k1) {
            const uint32_t key = uni32(rightmost_le(rp_in, rows, keys - 1, k));  // the key holding in-edge k
            const uint64_t re = (uint64_t)rp_in[key + 1];
            const uint64_t kend = re < k1 ? re : k1;
            const uint32_t slab = key - rows;
            const uint64_t wb = (uint64_t)slab * kSlabSpan;  // window start in colInd
            // the rows this segment names that the table lacks
            __syncthreads();  // the previous segment's lookups are done
            if (slab != cur)
                for (uint32_t i = threadIdx.x; i < kRefWords; i += kThreads) s_ref[i] = 0u;
            if (threadIdx.x == 0) s_nnew = 0u;
            cur = slab;
            __syncthreads();
            for (uint64_t x = k + threadIdx.x; x < kend; x += kThreads) {
                const uint32_t aux = auxG[x], st = aux & kSlabAuxLong, len = aux >> 13;
                if (len == kSlabAuxLong || len == 0) continue;  // searched in HBM, or empty
                const uint32_t bit = 1u << (st & 31u);
                if ((atomicOr(&s_ref[st >> 5], bit) & bit) == 0u) s_new[atomicAdd(&s_nnew, 1u)] = aux;
            }
            __syncthreads();
            // a wave per new row: its buckets cleared, then its ids inserted
            // (one wave's LDS operations stay in order)
            const uint32_t nnew = s_nnew;
            for (uint32_t r = wv; r < nnew; r += kWaves) {
                const uint32_t aux = s_new[r], st = aux & kSlabAuxLong, len = aux >> 13;
                for (uint32_t i = lane; i < len; i += kWave)
                    reinterpret_cast<uint4 *>(s_tab)[st + i] = make_uint4(kSentinel, kSentinel, kSentinel, kSentinel);
                for (uint32_t i = lane; i < len; i += kWave) {
                    const uint32_t u = colInd[wb + st + i];
                    uint32_t bk = st + slab_bucket(u, len);
                    for (bool done = false; !done;) {
                        for (uint32_t z = 0; z < kSlabWays && !done; z++)
                            done = atomicCAS(&s_tab[bk * kSlabWays + z], kSentinel, u) == kSentinel;
                        bk = bk + 1 == st + len ? st : bk + 1;
                    }
                }
            }
            __syncthreads();
            // the segment's in-edges dealt to the waves, kLanes lanes each
            constexpr uint32_t kPer = kWave / kLanes;
            const uint32_t sub = lane / kLanes, sl = lane % kLanes;
            uint32_t h = 0;
            for (uint64_t x0 = k + (uint64_t)wv * kPer; x0 < kend; x0 += (uint64_t)kWaves * kPer) {
                const uint64_t x = x0 + sub;
                OffT j = 0, hi = 0;
                uint32_t aux = 0;
                if (x < kend) {
                    j = ieG[2 * x];
                    hi = ieG[2 * x + 1];
                    aux = auxG[x];
                }
                OffT lo = j + 1;
                if (clipped && j == clip_j) lo = clip_lo;
                if (x >= kend) hi = lo;  // nothing to stream
                if (sl == 0) checked += (uint64_t)(hi - lo);
                const uint32_t st = aux & kSlabAuxLong, len = aux >> 13;
                if (len == 0) continue;  // adj+(w) is empty: no triangle
                OffT gb = 0, ge = 0;  // adj+(w) in HBM for a long row
                if (len == kSlabAuxLong) {
                    const uint32_t w = colInd[j];
                    gb = rowPtr[w];
                    ge = rowPtr[w + 1];
                }
                for (OffT q = (lo & ~(OffT)3) + 4 * sl; q < hi; q += 8 * kLanes) {  // two loads in flight
                    const OffT q1 = q + 4 * kLanes;
                    const u32x4 a = *reinterpret_cast<gptr<u32x4>>(colInd + q);
                    u32x4 b = {0u, 0u, 0u, 0u};
                    if (q1 < hi) b = *reinterpret_cast<gptr<u32x4>>(colInd + q1);
                    const uint32_t ua[8] = {a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
                    if (len != kSlabAuxLong) {
                        // each id's first bucket read for all eight at once (one 16-byte LDS
                        // read each; a position outside the tail reads the row's first
                        // bucket and is not counted), the four slots compared without
                        // branches; slots fill in order, so a free last slot ends the chain,
                        // and only a full bucket without the id goes on to the next
                        const u32x4 *tab4 = reinterpret_cast<const u32x4 *>(s_tab);
                        uint32_t bk[8];
                        bool in[8];
                        u32x4 v[8];
#pragma unroll
                        for (int z = 0; z < 8; z++) {
                            const OffT t = (z < 4 ? q : q1) + (OffT)(z & 3);
                            in[z] = t >= lo && t < hi;
                            bk[z] = st + slab_bucket(ua[z], len);
                            v[z] = tab4[in[z] ? bk[z] : st];
                        }
#pragma unroll
                        for (int z = 0; z < 8; z++) {
                            const uint32_t u = ua[z];
                            const bool f = (v[z][0] == u) | (v[z][1] == u) | (v[z][2] == u) | (v[z][3] == u);
                            h += (in[z] & f) ? 1u : 0u;
                            if (in[z] & !f & (v[z][3] != kSentinel)) {
                                uint32_t b2 = bk[z] + 1 == st + len ? st : bk[z] + 1;
                                for (;;) {
                                    const u32x4 w2 = tab4[b2];
                                    const bool f2 = (w2[0] == u) | (w2[1] == u) | (w2[2] == u) | (w2[3] == u);
                                    h += f2 ? 1u : 0u;
                                    if (f2 | (w2[3] == kSentinel)) break;
                                    b2 = b2 + 1 == st + len ? st : b2 + 1;
                                }
                            }
                        }
                        continue;
                    }
                    // a long row: adj+(w) searched in HBM
#pragma unroll
                    for (int z = 0; z < 8; z++) {
                        const OffT t = (z < 4 ? q : q1) + (OffT)(z & 3);
                        if (t < lo || t >= hi) continue;
                        const uint32_t u = ua[z];
                        OffT lb = gb, n_ = ge - gb;
                        while (n_ > 0) {
                            const OffT half = n_ >> 1;
                            if (colInd[lb + half] < u) {
                                lb += half + 1;
                                n_ -= half + 1;
                            } else {
                                n_ = half;
                            }
                        }
                        gb = lb;  // the lane's next id is larger
                        h += lb < ge && colInd[lb] == u ? 1u : 0u;
                    }
                }
            }
            hits += h;
            k = kend;
        }
    }
    hits = wave_sum(hits);
    checked = wave_sum(checked);
    if (lane == 0 && hits) atomicAdd(&counts[0], (unsigned long long)hits);
    if (lane == 0 && checked) atomicAdd(reinterpret_cast<unsigned long long *>(meta + kMetaListWedges), checked);
}

// ---------------------------------------------------------------- launches

hipError_t launch_split_prefix(const GraphDesc *desc, const SplitDesc *sd, int ob, uint32_t nblocks,
                               uint64_t *partials, hipStream_t st) {
    if (nblocks == 0) return hipSuccess;
    uint64_t *pf = partials, *pi = partials + nblocks;
    hipError_t e = WTC_LAUNCH_OFF(ob, k_split_partials, dim3(nblocks), dim3(kBlock), 0, st, desc, pf, pi);
    if (e != hipSuccess) return e;
    if ((e = launch_scan_partials(pf, nblocks, st)) != hipSuccess) return e;
    if ((e = launch_scan_partials(pi, nblocks, st)) != hipSuccess) return e;
    return WTC_LAUNCH_OFF(ob, k_split_scan, dim3(nblocks), dim3(kBlock), 0, st, desc, sd, (const uint64_t *)pf,
                          (const uint64_t *)pi);
}

// 8-byte row offsets that all fit 32 bits, copied to 4 bytes.
__global__ __launch_bounds__(kBlock) void k_split_narrow_offsets(const uint64_t *__restrict__ src,
                                                                uint32_t *__restrict__ dst, uint64_t count) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = (uint32_t)src[i];
}

hipError_t launch_split_narrow_offsets(const uint64_t *src, uint32_t *dst, uint64_t count, hipStream_t st) {
    if (count == 0) return hipSuccess;
    const uint64_t blocks = (count + kBlock - 1) / kBlock;
    k_split_narrow_offsets<<<dim3((uint32_t)(blocks < 4096 ? blocks : 4096)), dim3(kBlock), 0, st>>>(src, dst, count);
    return hipGetLastError();
}

hipError_t launch_split_range(const GraphDesc *desc, SplitDesc *sd, int ob, hipStream_t st) {
    return WTC_LAUNCH_OFF(ob, k_split_range, dim3(1), dim3(kSplitPartsMax), 0, st, desc, sd);
}

hipError_t launch_split_transpose(const GraphDesc *desc, const SplitDesc *sd, int ob, uint32_t chunks,
                                  uint32_t c2cap, uint32_t c3cap, uint32_t parts, uint32_t keys, uint64_t *key_part,
                                  uint64_t *meta, unsigned long long *counts, hipStream_t st) {
    if (keys == 0 || chunks == 0 || c2cap == 0 || c3cap == 0 || parts == 0 || parts > kSplitPartsMax)
        return hipErrorInvalidValue;
    const dim3 blk(kSplitBlock), g1(chunks);
    const dim3 g2(c2cap < kLevelGridMax ? c2cap : kLevelGridMax), g3(c3cap < kLevelGridMax ? c3cap : kLevelGridMax);
    const dim3 gk((keys + kBlock - 1) / kBlock);
    hipError_t e;
#define WTC_SPLIT_LAUNCH(KERNEL, GRID, BLOCK, ...)                                             \
    do {                                                                                       \
        if (ob == 8) KERNEL<uint64_t, __VA_ARGS__><<<GRID, BLOCK, 0, st>>>(sd);                \
        else KERNEL<uint32_t, __VA_ARGS__><<<GRID, BLOCK, 0, st>>>(sd);                        \
        if ((e = hipGetLastError()) != hipSuccess) return e;                                   \
    } while (0)
    if (ob == 8) k_split_l1count<uint64_t><<<g1, blk, 0, st>>>(desc, sd);
    else k_split_l1count<uint32_t><<<g1, blk, 0, st>>>(desc, sd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_split_l1cols<<<dim3(parts), blk, 0, st>>>(sd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_split_l1bases<<<dim3(1), dim3(kBlock), 0, st>>>(sd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ob == 8) k_split_l1<uint64_t><<<g1, blk, 0, st>>>(desc, sd);
    else k_split_l1<uint32_t><<<g1, blk, 0, st>>>(desc, sd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    WTC_SPLIT_LAUNCH(k_split_level_count, g2, blk, 2);
    k_split_levelcols2<<<dim3(parts), blk, 0, st>>>(sd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_split_subscan<<<dim3(1), blk, 0, st>>>(sd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    WTC_SPLIT_LAUNCH(k_split_level_scatter, g2, blk, 2);
    WTC_SPLIT_LAUNCH(k_split_level_count, g3, blk, 3);
    k_split_levelcols<3><<<gk, dim3(kBlock), 0, st>>>(sd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const dim3 gkb((keys + kKeyChunk - 1) / kKeyChunk);
    k_split_keysum<<<gkb, blk, 0, st>>>(sd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = launch_scan_partials(key_part, gkb.x, st)) != hipSuccess) return e;
    if (ob == 8) k_split_keyscan<uint64_t><<<gkb, blk, 0, st>>>(sd, meta, counts);
    else k_split_keyscan<uint32_t><<<gkb, blk, 0, st>>>(sd, meta, counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    WTC_SPLIT_LAUNCH(k_split_level_scatter, g3, blk, 3);
#undef WTC_SPLIT_LAUNCH
    // the packed top ids, over B's pairs (the third level has read them)
    return WTC_LAUNCH_OFF(ob, k_split_pack, dim3(kPackGrid), dim3(kBlock), 0, st, desc, sd);
}

size_t split_rows_lds(uint32_t rows) { return (size_t)(((rows + 31) / 32 + 1) & ~1u) * sizeof(uint32_t) + sizeof(uint64_t); }

template <int W, int L>
static const void *split_rows_fn(int ob) {
    return ob == 8 ? reinterpret_cast<const void *>(&k_split_rows<uint64_t, W, kSplitLanes, kSplitChunks, L>)
                   : reinterpret_cast<const void *>(&k_split_rows<uint32_t, W, kSplitLanes, kSplitChunks, L>);
}

template <int L>
static const void *split_rows_kernel_l(int ob, int wpb) {
    switch (wpb) {
        case 4: return split_rows_fn<4, L>(ob);
        case 16: return split_rows_fn<16, L>(ob);
        default: return split_rows_fn<8, L>(ob);
    }
}

// The row kernel for a graph of `rows` top rows (the packing's layout).
static const void *split_rows_kernel(int ob, int wpb, uint32_t rows) {
    switch (pack_layout(rows)) {
        case 0: return split_rows_kernel_l<0>(ob, wpb);
        case 1: return split_rows_kernel_l<1>(ob, wpb);
        case 2: return split_rows_kernel_l<2>(ob, wpb);
        default: return split_rows_kernel_l<3>(ob, wpb);
    }
}

bool valid_split_waves(int wpb) { return wpb == 4 || wpb == 8 || wpb == 16; }

hipError_t split_rows_grid(int ob, uint32_t rows, int wpb, int num_cus, uint32_t *grid) {
    if (!valid_split_waves(wpb)) return hipErrorInvalidValue;
    const void *fn = split_rows_kernel(ob, wpb, rows);
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
    if (!valid_split_waves(wpb) || grid == 0 || rows > kPackRowsMax) return hipErrorInvalidValue;
    const void *fn = split_rows_kernel(ob, wpb, rows);
    void *args[] = {&desc, &sd, &meta, &counts};
    return hipLaunchKernel(fn, dim3(grid), dim3(wpb * kWave), args, split_rows_lds(rows), st);
}

// The slab kernel's block: 8 waves, 4 lanes per in-edge tail (list tails
// are short: ~50 ids on R-MAT-24; 2.53 ms there at 4 lanes, 2.71 at 2, 2.73
// at 8).
constexpr int kSlabWaves = 8;
constexpr int kSlabLanes = 4;

static const void *split_slabs_kernel(int ob) {
    return ob == 8 ? reinterpret_cast<const void *>(&k_split_slabs<uint64_t, kSlabWaves, kSlabLanes>)
                   : reinterpret_cast<const void *>(&k_split_slabs<uint32_t, kSlabWaves, kSlabLanes>);
}

hipError_t split_slabs_grid(int ob, int num_cus, uint32_t *grid) {
    int per_cu = 0;
    hipError_t e =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, split_slabs_kernel(ob), kSlabWaves * kWave, 0);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    *grid = (uint32_t)per_cu * (uint32_t)(num_cus > 0 ? num_cus : 1);
    return hipSuccess;
}

hipError_t launch_split_slabs(const GraphDesc *desc, const SplitDesc *sd, int ob, uint64_t *meta,
                              unsigned long long *counts, uint32_t grid, hipStream_t st) {
    if (grid == 0) return hipErrorInvalidValue;
    void *args[] = {&desc, &sd, &meta, &counts};
    return hipLaunchKernel(split_slabs_kernel(ob), dim3(grid), dim3(kSlabWaves * kWave), args, 0, st);
}

}  // namespace wtc
