// wtc_api.hip -- device side of the C ABI: contexts, the asynchronous
// graph-set pipeline, resident graphs, the reference-style tc_GPU drop-ins
// and the batched path.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "wtc_internal.h"

using namespace wtc;

// Many oriented CSRs in one device allocation: every graph keeps its own
// rowPtr (offsets relative to its own colInd) and colInd (local ids).
struct wtc_batch {
    wtc_ctx *ctx;
    uint32_t G;
    int ob;
    std::vector<uint32_t> n;
    std::vector<uint64_t> m, row_off, col_off;  // element offsets into rowPtr / colInd
    std::vector<uint64_t> W;                    // wedges per graph (for the autotuner)
    DevBuf rowPtr, colInd;
    // Replays: the batch's own stream and workspace (a captured graph keeps
    // their addresses) and the graph of the last (spread, a, tuning) counted.
    wtc::Slot slot;
    hipGraphExec_t exec = nullptr;
    wtc::SetPlan plan;
    std::vector<uint32_t> key_s, key_a;
    int key_ways = 0, key_wpb = 0, key_kernel = 0;
    uint32_t key_work = 0;
    uint32_t key_chunk = 0, key_small = 0;
};

namespace wtc {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }

int hip_fail(const char *what, hipError_t e) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return WTC_ERR_HIP;
}

hipError_t DevBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) {
        hipError_t e = hipFree(p);
        p = nullptr;
        cap = 0;
        if (e != hipSuccess) return e;
    }
    // kDevTailSlack readable bytes past the end: the counting kernel's probe
    // window may load up to 15 elements beyond the last colInd entry.
    const size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, want + kDevTailSlack);
    if (e == hipSuccess) cap = want;
    else p = nullptr;
    return e;
}

hipError_t PinBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) {
        (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    const size_t want = std::max<size_t>(bytes, 4096);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e == hipSuccess) cap = want;
    else p = nullptr;
    return e;
}

int Slot::create() {
    WTC_HIPC(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (auto &e : ev) WTC_HIPC(hipEventCreate(&e));
    WTC_HIPC(hipEventCreateWithFlags(&ev_ready, hipEventDisableTiming));
    return WTC_OK;
}

void Slot::release_device() {
    ws.release();
    partials.release();
    tree.release();
    bm.release();
    desc.release();
    out.release();
    sp.release();
}

void Slot::destroy() {
    if (stream) (void)hipStreamSynchronize(stream);
    release_device();
    h_desc.release();
    h_out.release();
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (ev_ready) (void)hipEventDestroy(ev_ready);
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    ev_ready = nullptr;
    for (auto &e : ev) e = nullptr;
}

uint64_t bm_words_for(uint32_t a) { return (uint64_t)a * (a / 32u); }

int validate_oriented_serial(const wtc_big_graph *g) {
    if (!g || !g->rowPtr || (g->numEdges && !g->colInd)) return WTC_ERR_ARG;
    return validate_oriented_rows(g->numVertices, g->numEdges, g->rowPtr, g->colInd, 0, g->numVertices);
}

int check_args(uint32_t spread, uint32_t a) {
    if (spread == 0) return WTC_ERR_ARG;
    if (a % 32 != 0) return WTC_ERR_ARG;
    return WTC_OK;
}

GraphDesc make_desc(const void *rowPtr, const uint32_t *colInd, uint32_t n, uint32_t spread, uint32_t a,
                    uint64_t w_begin, uint64_t w_end) {
    GraphDesc d{};
    d.rowPtr = rowPtr;
    d.colInd = colInd;
    d.n = n;
    d.spread = spread;
    d.a = a;
    d.bm_rows = a ? std::min(n, a) : 0u;
    d.thr = a ? std::max(n, a) - a : n;
    d.bm_words = bm_words_for(a);
    d.w_begin = w_begin;
    d.w_end = w_end;
    return d;
}

// Blocks of the small-graph kernel for a graph of W wedges (W < 2^31).
static uint32_t small_parts_of(uint64_t W) {
    const uint64_t p = (W + kSmallBlockWedges - 1) / kSmallBlockWedges;
    return p < 1 ? 1u : (uint32_t)(p > 65535 ? 65535 : p);
}

static int ways_index(int w) { return w == 1 ? 0 : w == 2 ? 1 : w == 4 ? 2 : 3; }
static int wpb_index(int v) { return v == 1 ? 0 : v == 2 ? 1 : v == 4 ? 2 : 3; }

namespace {

// Byte offsets of the pieces of a split count's workspace in one buffer
// (each on a 256-byte boundary), and the buffer's size.
struct SplitLayout {
    size_t wsfull, hist1, part_base, c2base, hist2, sub_base, c3base, hist3, sub_tot, list_bid, slab_row, key_tot;
    size_t part_tot, key_part;
    size_t recid, rec, recaux, recid2, rec2, recaux2, rpin, rp32, total;
    uint32_t chunks, slabs, keys, parts, subs, c2cap, c3cap;
    uint64_t cap;
    int ob;  // offset width of the split kernels (4 for a narrowed 8-byte CSR)
};

// Whether a split count of a graph with 8-byte offsets and m edges runs on a
// 4-byte copy of its row offsets (every position fits 32 bits): records and
// offsets of 8 bytes instead of 16 through the transposition.  On unless
// $WTC_SPLIT_NARROW is 0.
bool split_narrow(int ob, uint64_t m) {
    if (ob != 8 || m >= 0xffffffffull) return false;
    const char *e = std::getenv("WTC_SPLIT_NARROW");
    return !(e && e[0] == '0');
}

// Slabs of the list rows of a graph of m edges.
uint64_t split_slabs(uint64_t m) { return (m + kSlabSpan - 1) / kSlabSpan; }

SplitLayout split_plan(uint32_t n, uint64_t m, uint32_t rows, int ob_in) {
    SplitLayout L{};
    const bool narrow = split_narrow(ob_in, m);
    const int ob = narrow ? 4 : ob_in;
    L.ob = ob;
    L.chunks = std::min<uint32_t>(kSplitChunksMax, std::max<uint32_t>(1, n / kSplitChunkVerts));
    L.slabs = (uint32_t)split_slabs(m);
    L.keys = rows + L.slabs;
    L.parts = (L.keys + kSplitPart - 1) / kSplitPart;
    L.subs = L.parts * kSplitFan2;
    L.cap = std::max<uint64_t>(m, 1);  // an in-edge is an edge
    L.c2cap = (uint32_t)((L.cap + kSplitChunk2 - 1) / kSplitChunk2) + L.parts;
    L.c3cap = (uint32_t)((L.cap + kSplitChunk3 - 1) / kSplitChunk3) + L.subs;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = (off + bytes + 255) & ~(size_t)255;
        return at;
    };
    const size_t pair = 2 * (size_t)ob;
    L.wsfull = take(2 * ((size_t)n + 1) * sizeof(uint64_t));  // ws_full, then ie_pre
    L.hist1 = take((size_t)L.chunks * L.parts * sizeof(uint64_t));
    L.part_base = take(((size_t)L.parts + 1) * sizeof(uint64_t));
    L.c2base = take(((size_t)L.parts + 1) * sizeof(uint32_t));
    L.hist2 = take((size_t)L.c2cap * kSplitFan2 * sizeof(uint32_t));
    L.sub_base = take(((size_t)L.subs + 1) * sizeof(uint64_t));
    L.c3base = take(((size_t)L.subs + 1) * sizeof(uint32_t));
    L.hist3 = take((size_t)L.c3cap * kSplitSub * sizeof(uint32_t));
    L.sub_tot = take((size_t)L.subs * sizeof(uint32_t));
    L.list_bid = take((size_t)L.parts * sizeof(uint32_t));
    L.slab_row = take(((size_t)L.slabs + 1) * sizeof(uint32_t));
    L.part_tot = take((size_t)L.parts * sizeof(uint64_t));
    L.key_part = take(((size_t)L.keys / kKeyChunk + 2) * sizeof(uint64_t));
    L.key_tot = take((size_t)L.keys * sizeof(uint32_t));
    L.recid = take((size_t)L.cap * sizeof(uint16_t));
    L.rec = take((size_t)L.cap * pair);
    L.recaux = take((size_t)L.cap * sizeof(uint32_t));
    L.recid2 = take((size_t)L.cap * sizeof(uint8_t));
    L.rec2 = take((size_t)L.cap * pair);
    L.recaux2 = take((size_t)L.cap * sizeof(uint32_t));
    L.rpin = take(((size_t)L.keys + 1) * (size_t)ob);
    L.rp32 = narrow ? take(((size_t)n + 1) * sizeof(uint32_t)) : 0;
    L.total = off;
    return L;
}

}  // namespace

size_t split_workspace_bytes(uint32_t n, uint64_t m, uint32_t rows, int ob) { return split_plan(n, m, rows, ob).total; }

int split_layout(SplitDesc *sd, char *base, uint32_t n, uint64_t m, uint32_t rows, int ob, uint32_t **rp32) {
    const SplitLayout L = split_plan(n, m, rows, ob);
    if (rp32) *rp32 = L.rp32 ? reinterpret_cast<uint32_t *>(base + L.rp32) : nullptr;
    sd->rows = rows;
    sd->keys = L.keys;
    sd->slabs = L.slabs;
    sd->chunks = L.chunks;
    sd->parts = L.parts;
    sd->subs = L.subs;
    sd->cap = L.cap;
    sd->c2cap = L.c2cap;
    sd->c3cap = L.c3cap;
    sd->ws_full = reinterpret_cast<uint64_t *>(base + L.wsfull);
    sd->ie_pre = sd->ws_full + (size_t)n + 1;
    sd->hist1 = reinterpret_cast<uint64_t *>(base + L.hist1);
    sd->part_base = reinterpret_cast<uint64_t *>(base + L.part_base);
    sd->c2base = reinterpret_cast<uint32_t *>(base + L.c2base);
    sd->hist2 = reinterpret_cast<uint32_t *>(base + L.hist2);
    sd->sub_base = reinterpret_cast<uint64_t *>(base + L.sub_base);
    sd->c3base = reinterpret_cast<uint32_t *>(base + L.c3base);
    sd->hist3 = reinterpret_cast<uint32_t *>(base + L.hist3);
    sd->sub_tot = reinterpret_cast<uint32_t *>(base + L.sub_tot);
    sd->list_bid = reinterpret_cast<uint32_t *>(base + L.list_bid);
    sd->slab_row = reinterpret_cast<uint32_t *>(base + L.slab_row);
    sd->part_tot = reinterpret_cast<uint64_t *>(base + L.part_tot);
    sd->key_part = reinterpret_cast<uint64_t *>(base + L.key_part);
    sd->key_tot = reinterpret_cast<uint32_t *>(base + L.key_tot);
    sd->rec_id = reinterpret_cast<uint16_t *>(base + L.recid);
    sd->rec = base + L.rec;
    sd->rec_aux = reinterpret_cast<uint32_t *>(base + L.recaux);
    sd->rec_id2 = reinterpret_cast<uint8_t *>(base + L.recid2);
    sd->rec2 = base + L.rec2;
    sd->rec_aux2 = reinterpret_cast<uint32_t *>(base + L.recaux2);
    sd->rp_in = base + L.rpin;
    sd->ie = sd->rec;  // the third level writes the pairs over A's
    sd->ie_aux = sd->rec_aux;
    return L.ob;
}

// A bitmap length whose row does not fit one block's LDS (or a graph whose
// keys overflow the transposition's parts) is counted by the row walk
// against the bitmap in HBM, as the reference does.
bool split_eligible(const Tuning &t, const GraphDesc &d, uint64_t m) {
    return t.kernel == kKernelSplit && d.mode == kModeTiles && d.a > 0 && d.bm_rows > 0 &&
           (uint64_t)d.bm_rows + split_slabs(m) <= (uint64_t)kSplitPartsMax * kSplitPart &&
           split_rows_lds(d.bm_rows) <= t.lds_bytes && (reinterpret_cast<uintptr_t>(d.colInd) & 15u) == 0;
}

int stage_set(Slot &s, Tuning &t, int ob, std::vector<GraphDesc> &hd, SetPlan *plan, const uint64_t *m,
              const uint64_t *W, char *split_base) {
    const uint32_t G = (uint32_t)hd.size();
    if (G == 0) return WTC_ERR_ARG;
    // Kernel choice per set: a set of two or more graphs that all fit the
    // small-graph limit is counted by the small-graph kernel (one block per
    // graph, CSR in LDS, so no bitmap); any other set by the wedge-parallel
    // kernel.  Mixing the two in one set runs them back to back, and the
    // small kernel's per-graph tail then costs more than it saves.
    uint32_t small_words = 0, small_n = 0;
    uint64_t small_parts = 0;
    const uint32_t small_cap = std::min(t.small_words, t.small_words_max);
    bool all_small = m && W && G >= 2 && small_cap > 0;
    for (uint32_t g = 0; g < G && all_small; g++)
        all_small = hd[g].n > 0 && 2 * ((uint64_t)hd[g].n + 1) + m[g] <= small_cap;
    for (uint32_t g = 0; g < G; g++) {
        GraphDesc &d = hd[g];
        d.mode = kModeTiles;
        if (!all_small) continue;
        const uint64_t words = 2 * ((uint64_t)d.n + 1) + m[g];
        d.mode = kModeSmall;
        d.a = 0;
        d.bm_rows = 0;
        d.thr = d.n;
        d.bm_words = 0;
        d.bitmap = nullptr;
        small_n++;
        small_parts += small_parts_of(W[g]);
        small_words = std::max(small_words, (uint32_t)words);
    }
    // The split count takes a single graph with a bitmap length whose colInd
    // is 16-byte aligned (its slab kernel loads four ids at a time); its top
    // rows then need no bitmap in HBM.
    const bool split = G == 1 && m != nullptr && split_eligible(t, hd[0], m[0]) && (!t.split_needs_base || split_base);
    const uint32_t split_rows = split ? hd[0].bm_rows : 0;
    if (split) {
        hd[0].bm_rows = 0;
        hd[0].bm_words = 0;
        hd[0].bitmap = nullptr;
    }
    uint32_t max_spread = 1;
    uint64_t ws_n = 0, bm_words = 0, rows = 0, blocks = 0;
    for (const auto &d : hd) {
        if (d.n == 0 || d.spread == 0) return WTC_ERR_ARG;
        if (d.mode == kModeTiles) max_spread = std::max(max_spread, d.spread);
        if (!d.ws && !split) ws_n += (uint64_t)d.n + 1;  // the split count keeps its prefixes in its workspace
        if (d.bm_rows && !d.bitmap) bm_words += bm_words_for(d.a);
        rows += d.bm_rows;
        blocks += prefix_blocks(d.n);
    }
    if (blocks > 0x7fffffffULL) return WTC_ERR_RANGE;
    const bool single = G == 1;
    if (ws_n) WTC_HIPC(s.ws.ensure(ws_n * sizeof(uint64_t)));
    WTC_HIPC(s.partials.ensure((split ? 2 : 1) * blocks * sizeof(uint64_t)));
    if (single) WTC_HIPC(s.tree.ensure(kTreeSize * sizeof(uint64_t)));
    // descriptors, then the block list of the small-graph kernel
    if (small_parts > 0x7fffffffULL) return WTC_ERR_RANGE;
    const size_t list_bytes = (size_t)G * sizeof(GraphDesc) + (size_t)small_parts * sizeof(SmallPart);
    const size_t split_off = (list_bytes + 255) & ~(size_t)255;
    const size_t desc_bytes = split ? split_off + sizeof(SplitDesc) : list_bytes;
    WTC_HIPC(s.desc.ensure(desc_bytes));
    WTC_HIPC(s.out.ensure((size_t)(kMetaSlots + G) * sizeof(uint64_t)));
    if (bm_words) WTC_HIPC(s.bm.ensure(bm_words * sizeof(uint32_t)));
    WTC_HIPC(s.h_desc.ensure(desc_bytes));
    WTC_HIPC(s.h_out.ensure((size_t)(kMetaSlots + G) * sizeof(uint64_t)));
    uint64_t ws_off = 0, bm_off = 0, row_off = 0, blk_off = 0, clear_off = 0;
    for (auto &d : hd) {
        d.bm_clear_base = clear_off;
        clear_off += d.bm_rows ? bm_words_for(d.a) : 0;
        if (!d.ws && !split) {
            d.ws = s.ws.as<uint64_t>() + ws_off;
            ws_off += (uint64_t)d.n + 1;
        }
        if (d.bm_rows && !d.bitmap) {
            d.bitmap = s.bm.as<uint32_t>() + bm_off;
            bm_off += bm_words_for(d.a);
        }
        d.row_base = row_off;
        d.blk_base = blk_off;
        row_off += d.bm_rows;
        blk_off += prefix_blocks(d.n);
    }
    SplitDesc sd{};
    int split_ob = ob;
    const uint64_t *split_rp_src = nullptr;
    uint32_t *split_rp_dst = nullptr;
    if (split) {
        char *base = split_base;
        if (!base) {
            WTC_HIPC(s.sp.ensure(split_workspace_bytes(hd[0].n, m[0], split_rows, ob)));
            base = s.sp.as<char>();
        }
        uint32_t *rp32 = nullptr;
        split_ob = split_layout(&sd, base, hd[0].n, m[0], split_rows, ob, &rp32);
        if (rp32) {  // the split kernels read the 4-byte copy the count starts by making
            split_rp_src = static_cast<const uint64_t *>(hd[0].rowPtr);
            split_rp_dst = rp32;
            hd[0].rowPtr = rp32;
        }
        WTC_HIPC(split_rows_grid(split_ob, split_rows, t.split_wpb, t.num_cus, &s.rows_grid));
        WTC_HIPC(split_slabs_grid(split_ob, t.num_cus, &s.slabs_grid));
    }
    std::memcpy(s.h_desc.p, hd.data(), (size_t)G * sizeof(GraphDesc));
    if (split) std::memcpy(s.h_desc.as<char>() + split_off, &sd, sizeof(SplitDesc));
    SmallPart *h_parts = reinterpret_cast<SmallPart *>(s.h_desc.as<GraphDesc>() + G);
    for (uint32_t g = 0, k = 0; g < G; g++) {
        if (hd[g].mode != kModeSmall) continue;
        const uint32_t np = small_parts_of(W[g]);
        for (uint32_t p = 0; p < np; p++) h_parts[k++] = SmallPart{g, (uint16_t)p, (uint16_t)np};
    }

    const int oi = ob == 8 ? 1 : 0;
    const uint32_t staged = std::min<uint32_t>(max_spread, t.chunk);
    // a set the split count does not take uses the row walk
    const int kernel = t.kernel == kKernelStaged ? kKernelStaged : kKernelWalk;
    uint32_t &gc = t.grid_cache[kernel == kKernelWalk ? 0 : 1][oi][ways_index(t.ways)][wpb_index(t.wpb)][staged];
    if (!gc) WTC_HIPC(count_grid(kernel, ob, staged, t.ways, t.wpb, t.num_cus, &gc));
    s.G = G;
    s.grid = gc;
    s.split = split;
    s.small = small_n;
    plan->G = G;
    plan->ob = ob;
    plan->blocks = (uint32_t)blocks;
    plan->rows = rows;
    plan->clear_words = clear_off;
    plan->max_spread = max_spread;
    plan->single = single;
    plan->ways = t.ways;
    plan->wpb = t.wpb;
    plan->kernel = kernel;
    plan->work = count_work(t.grab_for(kernel), t.deal, max_spread);
    plan->chunk = t.chunk;
    plan->desc_bytes = desc_bytes;
    plan->small_count = (uint32_t)small_parts;
    plan->tile_graphs = G - small_n;
    plan->small_lds = std::min((small_words * 4u + 255u) & ~255u, t.small_words_max * 4u);
    plan->split = split;
    plan->split_off = split_off;
    plan->split_rows = split_rows;
    plan->split_chunks = sd.chunks;
    plan->split_keys = sd.keys;
    plan->split_c2cap = sd.c2cap;
    plan->split_c3cap = sd.c3cap;
    plan->split_parts = sd.parts;
    plan->split_key_part = sd.key_part;
    plan->split_wpb = t.split_wpb;
    plan->split_ob = split_ob;
    plan->split_rp_src = split_rp_src;
    plan->split_rp_dst = split_rp_dst;
    plan->split_rp_n = (uint64_t)hd[0].n + 1;
    return WTC_OK;
}

int launch_set(Slot &s, const SetPlan &p, bool timed) {
    uint64_t *meta = s.out.as<uint64_t>();
    unsigned long long *dcounts = reinterpret_cast<unsigned long long *>(meta + kMetaSlots);
    GraphDesc *ddesc = s.desc.as<GraphDesc>();
    uint64_t *tree = p.single ? s.tree.as<uint64_t>() : nullptr;
    if (p.split) {
        SplitDesc *sd = reinterpret_cast<SplitDesc *>(s.desc.as<char>() + p.split_off);
        const int ob = p.split_ob;
        if (timed) WTC_HIPC(hipEventRecord(s.ev[0], s.stream));
        if (p.split_rp_dst) WTC_HIPC(launch_split_narrow_offsets(p.split_rp_src, p.split_rp_dst, p.split_rp_n, s.stream));
        WTC_HIPC(launch_split_prefix(ddesc, sd, ob, p.blocks, s.partials.as<uint64_t>(), s.stream));
        WTC_HIPC(launch_split_range(ddesc, sd, ob, s.stream));
        if (timed) WTC_HIPC(hipEventRecord(s.ev[1], s.stream));
        WTC_HIPC(launch_split_transpose(ddesc, sd, ob, p.split_chunks, p.split_c2cap, p.split_c3cap, p.split_parts,
                                        p.split_keys, p.split_key_part, meta, dcounts, s.stream));
        if (timed) WTC_HIPC(hipEventRecord(s.ev[2], s.stream));
        WTC_HIPC(launch_split_rows(ddesc, sd, ob, p.split_rows, p.split_wpb, meta, dcounts, s.rows_grid,
                                   s.stream));
        if (timed) WTC_HIPC(hipEventRecord(s.ev[4], s.stream));
        WTC_HIPC(launch_split_slabs(ddesc, sd, ob, meta, dcounts, s.slabs_grid, s.stream));
        if (timed) WTC_HIPC(hipEventRecord(s.ev[3], s.stream));
        WTC_HIPC(hipMemcpyAsync(s.h_out.p, meta, (size_t)(kMetaSlots + p.G) * sizeof(uint64_t),
                                hipMemcpyDeviceToHost, s.stream));
        return WTC_OK;
    }
    if (timed) WTC_HIPC(hipEventRecord(s.ev[0], s.stream));
    WTC_HIPC(launch_prefix(ddesc, p.G, p.ob, p.blocks, s.partials.as<uint64_t>(), s.stream));
    WTC_HIPC(launch_setup(ddesc, p.G, tree, meta, dcounts, s.stream));
    if (timed) WTC_HIPC(hipEventRecord(s.ev[1], s.stream));
    WTC_HIPC(launch_bitmap(ddesc, p.G, p.ob, p.rows, p.clear_words, s.stream));
    if (timed) WTC_HIPC(hipEventRecord(s.ev[2], s.stream));
    WTC_HIPC(launch_count_small(ddesc, reinterpret_cast<const SmallPart *>(ddesc + p.G), p.small_count, p.ob,
                                dcounts, p.small_lds, s.stream));
    if (p.tile_graphs)  // a set counted wholly by the small-graph kernel has no tiles
        WTC_HIPC(launch_count(p.kernel, ddesc, p.G, p.ob, tree, meta, dcounts, p.max_spread, p.chunk, p.ways, p.wpb,
                              p.work, s.grid,
                              s.stream));
    if (timed) WTC_HIPC(hipEventRecord(s.ev[3], s.stream));
    WTC_HIPC(hipMemcpyAsync(s.h_out.p, meta, (size_t)(kMetaSlots + p.G) * sizeof(uint64_t), hipMemcpyDeviceToHost,
                            s.stream));
    return WTC_OK;
}

int enqueue_set(Slot &s, Tuning &t, int ob, std::vector<GraphDesc> &hd, const uint64_t *m, const uint64_t *W,
                char *split_base) {
    SetPlan p;
    int rc = stage_set(s, t, ob, hd, &p, m, W, split_base);
    if (rc) return rc;
    WTC_HIPC(hipMemcpyAsync(s.desc.p, s.h_desc.p, p.desc_bytes, hipMemcpyHostToDevice,
                            s.stream));
    return launch_set(s, p, true);
}

uint32_t default_small_words() {
    const char *e = std::getenv("WTC_SMALL_GRAPH_WORDS");
    if (!e) return kSmallWordsDefault;
    const long v = std::atol(e);
    return v < 0 ? 0u : (uint32_t)std::min<long>(v, kSmallWordsMax);
}

int init_small_graphs(Tuning &t, const hipDeviceProp_t &prop, uint32_t want) {
    t.lds_bytes = prop.sharedMemPerBlock;
    t.small_words_max = (uint32_t)std::min<size_t>(prop.sharedMemPerBlock / 4, kSmallWordsMax);
    t.small_words = std::min(want, t.small_words_max);
    const hipError_t e = allow_small_lds(t.small_words_max * 4u);
    return e == hipSuccess ? WTC_OK : hip_fail("small-graph kernel LDS limit", e);
}

bool slot_done(Slot &s) { return hipStreamQuery(s.stream) == hipSuccess; }

int finish_set(Slot &s, uint64_t *counts, uint64_t *wedges, RunTimes *rt) {
    WTC_HIPC(hipStreamSynchronize(s.stream));
    const uint64_t *h = s.h_out.as<uint64_t>();
    for (uint32_t g = 0; g < s.G; g++) counts[g] = h[kMetaSlots + g];
    if (wedges) *wedges = h[kMetaWedges];
    if (rt) {
        WTC_HIPC(hipEventElapsedTime(&rt->prefix_ms, s.ev[0], s.ev[1]));
        WTC_HIPC(hipEventElapsedTime(&rt->bitmap_ms, s.ev[1], s.ev[2]));
        WTC_HIPC(hipEventElapsedTime(&rt->count_ms, s.ev[2], s.ev[3]));
        rt->split = s.split;
        if (s.split) {
            WTC_HIPC(hipEventElapsedTime(&rt->rows_ms, s.ev[2], s.ev[4]));
            WTC_HIPC(hipEventElapsedTime(&rt->list_ms, s.ev[4], s.ev[3]));
            rt->top_wedges = h[kMetaTopWedges];
            rt->in_edges = h[kMetaInEdges];
            rt->list_in_edges = h[kMetaAllEdges] - h[kMetaInEdges];
        }
    }
    return WTC_OK;
}

}  // namespace wtc

namespace {

bool valid_ways(int w) { return w == 1 || w == 2 || w == 4 || w == 8; }
bool valid_waves(int v) { return valid_count_waves(v); }

// Default probes per lane for new contexts: WTC_PROBE_WAYS (1, 2, 4 or 8),
// else 4.
int default_probe_ways() {
    const char *e = std::getenv("WTC_PROBE_WAYS");
    const int v = e ? std::atoi(e) : 4;
    return valid_ways(v) ? v : 4;
}

// Default waves per counting block for new contexts: WTC_COUNT_WAVES (1, 2,
// 4 or 8), else 4.
int default_count_waves() {
    const char *e = std::getenv("WTC_COUNT_WAVES");
    const int v = e ? std::atoi(e) : 4;
    return valid_waves(v) ? v : 4;
}

// Default counting kernel for new contexts: WTC_COUNT_KERNEL (0 row walk,
// 1 staged, 2 split), else the row walk.
int default_count_kernel() {
    const char *e = std::getenv("WTC_COUNT_KERNEL");
    const int v = e ? std::atoi(e) : kKernelWalk;
    return v == kKernelStaged || v == kKernelSplit ? v : kKernelWalk;
}

int set_device(int dev) {
    int cnt = 0;
    hipError_t e = hipGetDeviceCount(&cnt);
    if (e != hipSuccess || cnt <= 0) {
        set_error("no HIP device");
        return WTC_ERR_NODEV;
    }
    if (dev < 0 || dev >= cnt) return WTC_ERR_ARG;
    WTC_HIPC(hipSetDevice(dev));
    return WTC_OK;
}

// Synchronous count of a graph set on the context's slot.
int run_set(wtc_ctx *c, int ob, std::vector<GraphDesc> &hd, uint64_t *counts, wtc_count_stats *st,
            RunTimes *rt_out = nullptr, const uint64_t *m = nullptr, const uint64_t *wedges = nullptr) {
    auto t0 = std::chrono::steady_clock::now();
    int rc = enqueue_set(c->slot, c->tune, ob, hd, m, wedges);
    if (rc) return rc;
    RunTimes rt;
    uint64_t W = 0;
    rc = finish_set(c->slot, counts, &W, &rt);
    auto t1 = std::chrono::steady_clock::now();
    if (rc) return rc;
    if (rt_out) *rt_out = rt;
    if (st) {
        *st = wtc_count_stats{};
        st->prefix_ms = rt.prefix_ms;
        st->bitmap_ms = rt.bitmap_ms;
        st->count_ms = rt.count_ms;
        st->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->wedges = W;
        st->blocks = c->slot.grid;
        st->launches = 1;
        st->small_graphs = c->slot.small;
        st->split = rt.split ? 1u : 0u;
        st->rows_ms = rt.rows_ms;
        st->top_wedges = rt.top_wedges;
        st->in_edges = rt.in_edges;
        st->list_ms = rt.list_ms;
        st->list_in_edges = rt.list_in_edges;
    }
    return WTC_OK;
}

}  // namespace

namespace wtc {

// Process-wide context for the context-free entry points (tc_GPU has no
// context argument); serialised by a mutex like the reference's globals.
namespace {
std::mutex g_default_mu;
wtc_ctx *g_default = nullptr;
}  // namespace

std::mutex &default_ctx_mutex() { return g_default_mu; }

int default_context(wtc_ctx **out) {
    if (!g_default) {
        int rc = wtc_ctx_create(0, &g_default);
        if (rc) return rc;
        g_default->tune.auto_shape = std::getenv("WTC_COUNT_KERNEL") == nullptr;
    }
    *out = g_default;
    return WTC_OK;
}

}  // namespace wtc

namespace {

// A tuned shape's kernel, block waves, probe ways and row-kernel waves on a
// context.
int apply_tuned(wtc_ctx *c, const wtc_tuned &t) {
    if (!valid_ways(t.probe_ways) || !valid_waves(t.waves_per_block) || !valid_split_waves(t.split_waves) ||
        (t.kernel != WTC_KERNEL_WALK && t.kernel != WTC_KERNEL_STAGED && t.kernel != WTC_KERNEL_SPLIT))
        return WTC_ERR_ARG;
    c->tune.kernel = t.kernel == WTC_KERNEL_WALK ? kKernelWalk : t.kernel == WTC_KERNEL_STAGED ? kKernelStaged
                                                                                              : kKernelSplit;
    c->tune.ways = t.probe_ways;
    c->tune.wpb = t.waves_per_block;
    c->tune.split_wpb = t.split_waves;
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
    // the whole CSR is checked on the host before anything reaches the
    // device: an id out of range or not above its row, an unsorted row or a
    // decreasing offset would otherwise become out-of-bounds device accesses
    if ((rc = validate_oriented_par<OffT>(n, m, rowPtr, colInd))) return rc;
    std::lock_guard<std::mutex> lk(default_ctx_mutex());
    wtc_ctx *c;
    if ((rc = default_context(&c))) return rc;
    WTC_HIPC(hipSetDevice(c->device));
    // the caller's s and a, with the autotuner's kernel and block shape for
    // the graph's size bucket (unless $WTC_COUNT_KERNEL fixed the kernel)
    if (c->tune.auto_shape) {
        wtc_tuned tuned;
        if (wtc_autotune_shape(n, m, wedge_total_par<OffT>(n, rowPtr), 0, &tuned) == WTC_OK &&
            (rc = apply_tuned(c, tuned)))
            return rc;
    }
    hipStream_t st = c->slot.stream;
    DevBuf drow, dcol;
    auto cleanup = [&] {  // every device buffer is freed at the end of the call, like the reference
        c->slot.release_device();
        drow.release();
        dcol.release();
    };
    WTC_HIPC(hipEventRecord(c->slot.ev[0], st));
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
    if ((e = hipEventRecord(c->slot.ev[1], st)) != hipSuccess ||
        (e = hipEventSynchronize(c->slot.ev[1])) != hipSuccess) {
        cleanup();
        return hip_fail("tc_gpu copy sync", e);
    }
    float copy_ms = 0;
    (void)hipEventElapsedTime(&copy_ms, c->slot.ev[0], c->slot.ev[1]);
    std::vector<GraphDesc> hd{make_desc(drow.p, dcol.as<uint32_t>(), n, spread, a, 0, UINT64_MAX)};
    uint64_t cnt = 0;
    RunTimes rt;
    const uint64_t m64 = m;
    rc = run_set(c, (int)sizeof(OffT), hd, &cnt, nullptr, &rt, &m64);
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
    c->tune.num_cus = prop.multiProcessorCount;
    c->tune.ways = default_probe_ways();
    c->tune.wpb = default_count_waves();
    c->tune.kernel = default_count_kernel();
    if ((rc = init_small_graphs(c->tune, prop, default_small_words()))) {
        delete c;
        return rc;
    }
    if ((rc = c->slot.create())) {
        wtc_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return WTC_OK;
}

int wtc_ctx_set_small_graphs(wtc_ctx *c, uint32_t max_words) {
    if (!c || max_words > c->tune.small_words_max) return WTC_ERR_ARG;
    c->tune.small_words = max_words;
    return WTC_OK;
}

int wtc_ctx_get_small_graphs(const wtc_ctx *c, uint32_t *max_words, uint32_t *device_max_words) {
    if (!c) return WTC_ERR_ARG;
    if (max_words) *max_words = c->tune.small_words;
    if (device_max_words) *device_max_words = c->tune.small_words_max;
    return WTC_OK;
}

int wtc_ctx_set_tuning(wtc_ctx *c, int probe_ways, uint32_t lds_chunk, int waves_per_block) {
    if (!c || !valid_ways(probe_ways) || lds_chunk < 1 || lds_chunk > kChunkMax || !valid_waves(waves_per_block))
        return WTC_ERR_ARG;
    c->tune.ways = probe_ways;
    c->tune.chunk = lds_chunk;
    c->tune.wpb = waves_per_block;
    return WTC_OK;
}

int wtc_ctx_set_work_split(wtc_ctx *c, uint32_t tiles_per_grab, int deal_to_xcds) {
    if (!c || tiles_per_grab > 0xffffu) return WTC_ERR_ARG;
    c->tune.grab = tiles_per_grab;
    c->tune.deal = deal_to_xcds != 0;
    return WTC_OK;
}

int wtc_ctx_set_count_kernel(wtc_ctx *c, int kernel) {
    if (!c || (kernel != WTC_KERNEL_WALK && kernel != WTC_KERNEL_STAGED && kernel != WTC_KERNEL_SPLIT))
        return WTC_ERR_ARG;
    c->tune.kernel = kernel == WTC_KERNEL_WALK ? kKernelWalk : kernel == WTC_KERNEL_STAGED ? kKernelStaged
                                                                                          : kKernelSplit;
    return WTC_OK;
}

int wtc_ctx_use_tuned(wtc_ctx *c, const wtc_tuned *t) {
    if (!c || !t) return WTC_ERR_ARG;
    return apply_tuned(c, *t);
}

void wtc_ctx_destroy(wtc_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    c->slot.destroy();
    for (auto &p : c->pipe) p.destroy();
    c->h_stage.release();
    c->batch_row.release();
    c->batch_col.release();
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
    WTC_HIPC(hipSetDevice(c->device));
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

int wtc_dgraph_info(const wtc_dgraph *dg, uint32_t *n, uint64_t *m, uint32_t *offset_bytes) {
    if (!dg) return WTC_ERR_ARG;
    if (n) *n = dg->n;
    if (m) *m = dg->m;
    if (offset_bytes) *offset_bytes = (uint32_t)dg->ob;
    return WTC_OK;
}

uint64_t wtc_dgraph_bytes(const wtc_dgraph *dg) {
    return dg ? ((uint64_t)dg->n + 1) * dg->ob + dg->m * 4u : 0;
}

int wtc_dgraph_count(wtc_ctx *c, wtc_dgraph *dg, uint32_t spread, uint32_t a, uint64_t w_begin, uint64_t w_end,
                     uint64_t *triangles, wtc_count_stats *st) {
    if (!c || !dg || !triangles) return WTC_ERR_ARG;
    int rc = check_args(spread, a);
    if (rc) return rc;
    WTC_HIPC(hipSetDevice(c->device));
    std::vector<GraphDesc> hd{make_desc(dg->rowPtr.p, dg->colInd.as<uint32_t>(), dg->n, spread, a, w_begin, w_end)};
    return run_set(c, dg->ob, hd, triangles, st, nullptr, &dg->m);
}

// Copies the wedge prefix and search tree left by the last single-graph count
// on this context (diagnostics for the parity tests).
int wtc_ctx_debug_prefix(wtc_ctx *c, uint64_t n, uint64_t *wedge_sum, uint64_t *tree) {
    if (!c) return WTC_ERR_ARG;
    WTC_HIPC(hipSetDevice(c->device));
    if (wedge_sum) {
        if (c->slot.ws.cap < (n + 1) * sizeof(uint64_t)) return WTC_ERR_ARG;
        WTC_HIPC(hipMemcpy(wedge_sum, c->slot.ws.p, (n + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
    }
    if (tree) {
        if (!c->slot.tree.p) return WTC_ERR_ARG;
        WTC_HIPC(hipMemcpy(tree, c->slot.tree.p, kTreeSize * sizeof(uint64_t), hipMemcpyDeviceToHost));
    }
    return WTC_OK;
}

int wtc_ctx_debug_bitmap(wtc_ctx *c, uint64_t words, uint32_t *out) {
    if (!c || !out) return WTC_ERR_ARG;
    if (c->slot.bm.cap < words * sizeof(uint32_t)) return WTC_ERR_ARG;
    WTC_HIPC(hipSetDevice(c->device));
    WTC_HIPC(hipMemcpy(out, c->slot.bm.p, words * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return WTC_OK;
}

int wtc_debug_decode(wtc_ctx *c, const uint32_t *d, const uint64_t *iv, uint64_t k, uint32_t *w_out,
                     uint32_t *u_out) {
    if (!c || (k && (!d || !iv || !w_out || !u_out))) return WTC_ERR_ARG;
    for (uint64_t i = 0; i < k; i++)
        if (d[i] < 2 || iv[i] >= binom2(d[i])) return WTC_ERR_ARG;
    if (k == 0) return WTC_OK;
    WTC_HIPC(hipSetDevice(c->device));
    hipStream_t st = c->slot.stream;
    DevBuf b;
    hipError_t e = b.ensure(k * (4 + 8 + 4 + 4));
    if (e != hipSuccess) return hip_fail("debug_decode alloc", e);
    char *p = b.as<char>();
    uint64_t *di = reinterpret_cast<uint64_t *>(p);  // 8-byte entries first, at the aligned base
    uint32_t *dd = reinterpret_cast<uint32_t *>(p + 8 * k), *dw = dd + k, *du = dw + k;
    if ((e = hipMemcpyAsync(dd, d, k * 4, hipMemcpyHostToDevice, st)) == hipSuccess &&
        (e = hipMemcpyAsync(di, iv, k * 8, hipMemcpyHostToDevice, st)) == hipSuccess &&
        (e = launch_debug_decode(dd, di, k, dw, du, st)) == hipSuccess &&
        (e = hipMemcpyAsync(w_out, dw, k * 4, hipMemcpyDeviceToHost, st)) == hipSuccess &&
        (e = hipMemcpyAsync(u_out, du, k * 4, hipMemcpyDeviceToHost, st)) == hipSuccess)
        e = hipStreamSynchronize(st);
    b.release();
    return e == hipSuccess ? WTC_OK : hip_fail("debug_decode", e);
}

int wtc_debug_copy(wtc_ctx *c, uint64_t bytes, int reps, double *ms) {
    if (!c || bytes < 16 || reps < 1) return WTC_ERR_ARG;
    bytes &= ~(uint64_t)15;
    WTC_HIPC(hipSetDevice(c->device));
    hipStream_t st = c->slot.stream;
    DevBuf src, dst;
    hipError_t e = src.ensure(bytes);
    if (e == hipSuccess) e = dst.ensure(bytes);
    if (e == hipSuccess) e = launch_fill(src.as<uint32_t>(), bytes / 4, 0x5a5a5a5au, st);
    if (e == hipSuccess) e = hipEventRecord(c->slot.ev[0], st);
    for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_copy(src.p, dst.p, bytes, st);
    if (e == hipSuccess) e = hipEventRecord(c->slot.ev[1], st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, c->slot.ev[0], c->slot.ev[1]);
    src.release();
    dst.release();
    if (e != hipSuccess) return hip_fail("debug_copy", e);
    if (ms) *ms = t / reps;
    return WTC_OK;
}

// ------------------------------------------------------------------ batches

}  // extern "C"


namespace {

// Copies graph gr into pinned staging (row offsets at element ro in the
// set's width ob, colInd at element co), checks that it is an oriented CSR
// (rows strictly ascending, every id above its row and below n) and returns
// its wedge count: one pass over the graph's data.
int pack_graph(const wtc_big_graph *gr, int ob, char *row_stage, uint32_t *col_stage, uint64_t ro, uint64_t co,
               uint64_t *W) {
    if (!gr || !gr->rowPtr || (gr->numEdges && !gr->colInd)) return WTC_ERR_ARG;
    const uint64_t n = gr->numVertices, m = gr->numEdges;
    const uint64_t *rp = gr->rowPtr;
    const uint32_t *ci = gr->colInd;
    if (rp[0] != 0 || rp[n] != m) return WTC_ERR_CSR;
    if (m) std::memcpy(col_stage + co, ci, m * 4);
    uint32_t *r32 = ob == 4 ? reinterpret_cast<uint32_t *>(row_stage) + ro : nullptr;
    if (ob == 8) std::memcpy(reinterpret_cast<uint64_t *>(row_stage) + ro, rp, (n + 1) * 8);
    uint64_t w = 0;
    for (uint64_t v = 0; v < n; v++) {
        const uint64_t s = rp[v], t = rp[v + 1];
        if (t < s || t > m) return WTC_ERR_CSR;
        if (r32) r32[v] = (uint32_t)s;
        w += binom2(t - s);
        uint64_t prev = v;
        for (uint64_t j = s; j < t; j++) {
            const uint64_t x = ci[j];
            if (x <= prev || x >= n) return WTC_ERR_CSR;
            prev = x;
        }
    }
    if (r32) r32[n] = (uint32_t)m;
    *W = w;
    return WTC_OK;
}

// Bytes of packed graphs after which a host-to-device copy is started, so
// that packing the next graphs overlaps the copy of the previous ones:
// $WTC_UPLOAD_CHUNK_BYTES, else 256 MiB.
constexpr uint64_t kUploadChunkBytes = 256ull << 20;
uint64_t upload_chunk_bytes() {
    const char *e = std::getenv("WTC_UPLOAD_CHUNK_BYTES");
    const long long v = e ? std::atoll(e) : 0;
    return v > 0 ? (uint64_t)v : kUploadChunkBytes;
}

// Packs k host CSRs into b's device rowPtr/colInd (grown as needed): sizes
// and offsets, then one pass per graph (copy into the context's pinned
// staging + check + wedge count) in chunks whose copies to the device run
// while the next chunk is packed.
int batch_fill(wtc_ctx *c, const wtc_big_graph *const *graphs, size_t k, wtc_batch *b) {
    b->G = (uint32_t)k;
    b->n.resize(k);
    b->m.resize(k);
    b->row_off.resize(k);
    b->col_off.resize(k);
    b->W.resize(k);
    uint64_t R = 0, M = 0, max_m = 0;
    for (size_t g = 0; g < k; g++) {
        if (!graphs[g] || graphs[g]->numVertices == 0) return WTC_ERR_ARG;
        b->n[g] = graphs[g]->numVertices;
        b->m[g] = graphs[g]->numEdges;
        b->row_off[g] = R;
        b->col_off[g] = M;
        R += (uint64_t)graphs[g]->numVertices + 1;
        M += graphs[g]->numEdges;
        max_m = std::max<uint64_t>(max_m, graphs[g]->numEdges);
    }
    b->ob = max_m < 0xffffffffULL ? 4 : 8;
    const int ob = b->ob;
    WTC_HIPC(hipSetDevice(c->device));
    const size_t rbytes = R * (size_t)ob, cbytes = std::max<uint64_t>(M, 1) * 4;
    WTC_HIPC(b->rowPtr.ensure(rbytes));
    WTC_HIPC(b->colInd.ensure(cbytes));
    WTC_HIPC(c->h_stage.ensure(rbytes + cbytes));
    char *stage = c->h_stage.as<char>();
    uint32_t *sc = reinterpret_cast<uint32_t *>(stage + rbytes);
    hipStream_t st = c->slot.stream;
    std::atomic<int> bad{WTC_OK};
    hipError_t e = hipSuccess;
    const uint64_t chunk_bytes = upload_chunk_bytes();
    for (size_t g0 = 0; g0 < k && e == hipSuccess;) {
        size_t g1 = g0;
        uint64_t bytes = 0;
        while (g1 < k && (g1 == g0 || bytes < chunk_bytes)) {
            bytes += (b->n[g1] + 1) * (uint64_t)ob + b->m[g1] * 4;
            g1++;
        }
        par_chunks(g1 - g0, 1, [&](uint64_t i0, uint64_t i1, int) {
            for (uint64_t i = g0 + i0; i < g0 + i1; i++) {
                const int rc = pack_graph(graphs[i], ob, stage, sc, b->row_off[i], b->col_off[i], &b->W[i]);
                if (rc) bad = rc;
            }
        });
        const uint64_t r0 = b->row_off[g0], r1 = b->row_off[g1 - 1] + b->n[g1 - 1] + 1;
        const uint64_t c0 = b->col_off[g0], c1 = b->col_off[g1 - 1] + b->m[g1 - 1];
        e = hipMemcpyAsync(b->rowPtr.as<char>() + r0 * ob, stage + r0 * ob, (r1 - r0) * ob, hipMemcpyHostToDevice,
                           st);
        if (e == hipSuccess && c1 > c0)
            e = hipMemcpyAsync(b->colInd.as<uint32_t>() + c0, sc + c0, (c1 - c0) * 4, hipMemcpyHostToDevice, st);
        g0 = g1;
    }
    const hipError_t es = hipStreamSynchronize(st);  // the staging is reused by the next call
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) return hip_fail("batch upload", e);
    return bad.load();
}

}  // namespace

extern "C" {

int wtc_batch_upload(wtc_ctx *c, const wtc_big_graph *const *graphs, size_t k, wtc_batch **out) {
    if (!c || !out || (k && !graphs) || k > 0xffffffffULL) return WTC_ERR_ARG;
    *out = nullptr;
    wtc_batch *b = new (std::nothrow) wtc_batch();
    if (!b) return WTC_ERR_NOMEM;
    b->ctx = c;
    const int rc = batch_fill(c, graphs, k, b);
    if (rc) {
        wtc_batch_free(b);
        return rc;
    }
    *out = b;
    return WTC_OK;
}

void wtc_batch_free(wtc_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->slot.destroy();
    b->rowPtr.release();
    b->colInd.release();
    delete b;
}

}  // extern "C"

namespace {

// Per-graph spread and bitmap length of a batch count (autotuned where 0).
// Per-graph spread and bitmap length of graphs [g0, g1) of a batch count
// (autotuned where 0); s_out and a_out hold the whole batch.
int batch_params(const wtc_batch *b, const uint32_t *spread, const uint32_t *a, std::vector<uint32_t> &s_out,
                 std::vector<uint32_t> &a_out, uint32_t g0 = 0, uint32_t g1 = UINT32_MAX) {
    s_out.resize(b->G);
    a_out.resize(b->G);
    for (uint32_t g = g0; g < std::min(g1, b->G); g++) {
        uint32_t s = spread ? spread[g] : 0, aa = a ? a[g] : 0;
        if (s == 0) wtc_autotune_batch(b->n[g], b->m[g], b->W[g], &s, a ? nullptr : &aa);
        if (s == 0 || aa % 32 != 0) return WTC_ERR_ARG;
        s_out[g] = s;
        a_out[g] = aa;
    }
    return WTC_OK;
}

// Descriptors of graphs [g0, g1) of a batch (the whole batch by default).
std::vector<GraphDesc> batch_descs(const wtc_batch *b, const std::vector<uint32_t> &s,
                                   const std::vector<uint32_t> &a, uint32_t g0 = 0, uint32_t g1 = UINT32_MAX) {
    g1 = std::min(g1, b->G);
    std::vector<GraphDesc> hd(g1 > g0 ? g1 - g0 : 0);
    for (uint32_t g = g0; g < g1; g++) {
        const void *rp = b->rowPtr.as<char>() + b->row_off[g] * b->ob;
        hd[g - g0] = make_desc(rp, b->colInd.as<uint32_t>() + b->col_off[g], b->n[g], s[g], a[g], 0, UINT64_MAX);
    }
    return hd;
}

// Captures the batch's launch sequence (kernels + the copy of the counts to
// the host) into a graph on the batch's stream.  The descriptors go to the
// device before the capture, so the graph holds no host-to-device copy and
// no memset: every replay recomputes the prefix, tree, bitmap and counts
// from the resident CSRs.
int batch_capture(wtc_ctx *c, wtc_batch *b, const std::vector<uint32_t> &s, const std::vector<uint32_t> &a) {
    if (b->exec) {
        WTC_HIPC(hipGraphExecDestroy(b->exec));
        b->exec = nullptr;
    }
    if (!b->slot.stream) {
        int rc = b->slot.create();
        if (rc) return rc;
    }
    std::vector<GraphDesc> hd = batch_descs(b, s, a);
    int rc = stage_set(b->slot, c->tune, b->ob, hd, &b->plan, b->m.data(), b->W.data());
    if (rc) return rc;
    hipStream_t st = b->slot.stream;
    WTC_HIPC(hipMemcpyAsync(b->slot.desc.p, b->slot.h_desc.p, b->plan.desc_bytes,
                            hipMemcpyHostToDevice, st));
    WTC_HIPC(hipStreamSynchronize(st));
    WTC_HIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    rc = launch_set(b->slot, b->plan, false);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    if (rc == WTC_OK && ec != hipSuccess) rc = hip_fail("batch capture", ec);
    if (rc == WTC_OK) {
        const hipError_t ei = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            b->exec = nullptr;
            rc = hip_fail("batch graph instantiate", ei);
        }
    }
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) return rc;
    b->key_s = s;
    b->key_a = a;
    b->key_ways = c->tune.ways;
    b->key_wpb = c->tune.wpb;
    b->key_kernel = c->tune.kernel;
    b->key_work = count_work(c->tune.grab_for(c->tune.kernel), c->tune.deal);
    b->key_chunk = c->tune.chunk;
    b->key_small = c->tune.small_words;
    return WTC_OK;
}

// Host CSRs counted as they arrive: the graphs are packed into pinned
// staging and copied to the device in chunks of about upload_chunk_bytes() on
// the context's stream (the upload stream here), and every chunk is counted
// as its own graph set on one of two execution slots as soon as its copy has
// landed, so counting overlaps packing and copying the later chunks.  A
// graph that is not an oriented CSR fails the whole call.
int batch_count_pipelined(wtc_ctx *c, const wtc_big_graph *const *graphs, size_t k, const uint32_t *spread,
                          const uint32_t *a, wtc_batch *b, uint64_t *triangles, RunTimes *sum, uint64_t *wsum,
                          uint32_t *launches, uint32_t *small) {
    b->G = (uint32_t)k;
    b->n.resize(k);
    b->m.resize(k);
    b->row_off.resize(k);
    b->col_off.resize(k);
    b->W.resize(k);
    uint64_t R = 0, M = 0, max_m = 0;
    for (size_t g = 0; g < k; g++) {
        if (!graphs[g] || graphs[g]->numVertices == 0) return WTC_ERR_ARG;
        b->n[g] = graphs[g]->numVertices;
        b->m[g] = graphs[g]->numEdges;
        b->row_off[g] = R;
        b->col_off[g] = M;
        R += (uint64_t)graphs[g]->numVertices + 1;
        M += graphs[g]->numEdges;
        max_m = std::max<uint64_t>(max_m, graphs[g]->numEdges);
    }
    b->ob = max_m < 0xffffffffULL ? 4 : 8;
    const int ob = b->ob;
    WTC_HIPC(hipSetDevice(c->device));
    const size_t rbytes = R * (size_t)ob, cbytes = std::max<uint64_t>(M, 1) * 4;
    WTC_HIPC(b->rowPtr.ensure(rbytes));
    WTC_HIPC(b->colInd.ensure(cbytes));
    WTC_HIPC(c->h_stage.ensure(rbytes + cbytes));
    for (auto &p : c->pipe)
        if (!p.stream) {
            const int rc = p.create();
            if (rc) return rc;
        }
    char *stage = c->h_stage.as<char>();
    uint32_t *sc = reinterpret_cast<uint32_t *>(stage + rbytes);
    hipStream_t up = c->slot.stream;
    std::vector<uint32_t> s_all, a_all;
    uint32_t pending_g0[2] = {0, 0};
    bool pending[2] = {false, false};
    int rc = WTC_OK;
    // collects the counts of the set in flight on pipe slot i
    auto drain = [&](int i) -> int {
        if (!pending[i]) return WTC_OK;
        pending[i] = false;
        RunTimes rt;
        uint64_t W = 0;
        const int r = finish_set(c->pipe[i], triangles + pending_g0[i], &W, &rt);
        if (r) return r;
        sum->prefix_ms += rt.prefix_ms;
        sum->bitmap_ms += rt.bitmap_ms;
        sum->count_ms += rt.count_ms;
        *wsum += W;
        *small += c->pipe[i].small;
        return WTC_OK;
    };
    int next = 0;
    const uint64_t chunk_bytes = upload_chunk_bytes();
    for (size_t g0 = 0; g0 < k && rc == WTC_OK;) {
        size_t g1 = g0;
        uint64_t bytes = 0;
        while (g1 < k && (g1 == g0 || bytes < chunk_bytes)) {
            bytes += (b->n[g1] + 1) * (uint64_t)ob + b->m[g1] * 4;
            g1++;
        }
        std::atomic<int> bad{WTC_OK};
        par_chunks(g1 - g0, 1, [&](uint64_t i0, uint64_t i1, int) {
            for (uint64_t i = g0 + i0; i < g0 + i1; i++) {
                const int r = pack_graph(graphs[i], ob, stage, sc, b->row_off[i], b->col_off[i], &b->W[i]);
                if (r) bad = r;
            }
        });
        if ((rc = bad.load())) break;
        const uint64_t r0 = b->row_off[g0], r1 = b->row_off[g1 - 1] + b->n[g1 - 1] + 1;
        const uint64_t c0 = b->col_off[g0], c1 = b->col_off[g1 - 1] + b->m[g1 - 1];
        hipError_t e = hipMemcpyAsync(b->rowPtr.as<char>() + r0 * ob, stage + r0 * ob, (r1 - r0) * ob,
                                      hipMemcpyHostToDevice, up);
        if (e == hipSuccess && c1 > c0)
            e = hipMemcpyAsync(b->colInd.as<uint32_t>() + c0, sc + c0, (c1 - c0) * 4, hipMemcpyHostToDevice, up);
        Slot &x = c->pipe[next];
        if (e == hipSuccess) e = hipEventRecord(x.ev_ready, up);
        if (e != hipSuccess) {
            rc = hip_fail("batch upload", e);
            break;
        }
        if ((rc = drain(next))) break;  // the slot's previous set (two chunks back)
        if ((rc = batch_params(b, spread, a, s_all, a_all, (uint32_t)g0, (uint32_t)g1))) break;
        std::vector<GraphDesc> hd = batch_descs(b, s_all, a_all, (uint32_t)g0, (uint32_t)g1);
        if ((e = hipStreamWaitEvent(x.stream, x.ev_ready, 0)) != hipSuccess) {
            rc = hip_fail("batch upload wait", e);
            break;
        }
        if ((rc = enqueue_set(x, c->tune, ob, hd, b->m.data() + g0, b->W.data() + g0))) break;
        pending[next] = true;
        pending_g0[next] = (uint32_t)g0;
        ++*launches;
        next ^= 1;
        g0 = g1;
    }
    // every copy and launch has finished before the staging or buffers are reused
    (void)hipStreamSynchronize(up);
    for (int i = 0; i < 2; i++) {
        const int r = drain(i);
        if (rc == WTC_OK) rc = r;
        if (c->pipe[i].stream) (void)hipStreamSynchronize(c->pipe[i].stream);
    }
    return rc;
}

}  // namespace

extern "C" {

int wtc_batch_count(wtc_ctx *c, wtc_batch *b, const uint32_t *spread, const uint32_t *a, uint64_t *triangles,
                    wtc_count_stats *st) {
    if (!c || !b || !triangles || b->ctx != c) return WTC_ERR_ARG;
    if (b->G == 0) return WTC_OK;
    const auto t0 = std::chrono::steady_clock::now();
    WTC_HIPC(hipSetDevice(c->device));
    std::vector<uint32_t> s, aa;
    int rc = batch_params(b, spread, a, s, aa);
    if (rc) return rc;
    if (!b->exec || s != b->key_s || aa != b->key_a || c->tune.ways != b->key_ways || c->tune.wpb != b->key_wpb || c->tune.kernel != b->key_kernel ||
        count_work(c->tune.grab_for(c->tune.kernel), c->tune.deal) != b->key_work ||
        c->tune.chunk != b->key_chunk || c->tune.small_words != b->key_small) {
        if ((rc = batch_capture(c, b, s, aa))) return rc;
    }
    Slot &sl = b->slot;
    WTC_HIPC(hipEventRecord(sl.ev[0], sl.stream));
    WTC_HIPC(hipGraphLaunch(b->exec, sl.stream));
    WTC_HIPC(hipEventRecord(sl.ev[3], sl.stream));
    WTC_HIPC(hipStreamSynchronize(sl.stream));
    const uint64_t *h = sl.h_out.as<uint64_t>();
    for (uint32_t g = 0; g < b->G; g++) triangles[g] = h[kMetaSlots + g];
    if (st) {
        *st = wtc_count_stats{};
        float ms = 0;
        WTC_HIPC(hipEventElapsedTime(&ms, sl.ev[0], sl.ev[3]));
        st->prefix_ms = 0;
        st->bitmap_ms = 0;
        st->count_ms = ms;
        st->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        st->wedges = h[kMetaWedges];
        st->blocks = sl.grid;
        st->launches = 1;
        st->small_graphs = sl.small;
    }
    return WTC_OK;
}

int wtc_batch_count_host(wtc_ctx *c, const wtc_big_graph *const *graphs, size_t k, const uint32_t *spread,
                         const uint32_t *a, uint64_t *triangles, wtc_count_stats *st) {
    auto t0 = std::chrono::steady_clock::now();
    if (!c || !triangles || (k && !graphs) || k > 0xffffffffULL) return WTC_ERR_ARG;
    if (k == 0) {
        if (st) *st = wtc_count_stats{};
        return WTC_OK;
    }
    wtc_batch *b = new (std::nothrow) wtc_batch();
    if (!b) return WTC_ERR_NOMEM;
    b->ctx = c;
    // the graphs go into the context's grow-only device buffers, so repeated
    // calls allocate no device memory once the buffers are large enough
    b->rowPtr = std::move(c->batch_row);
    b->colInd = std::move(c->batch_col);
    RunTimes sum{};
    uint64_t wsum = 0;
    uint32_t launches = 0, small = 0;
    int rc = batch_count_pipelined(c, graphs, k, spread, a, b, triangles, &sum, &wsum, &launches, &small);
    c->batch_row = std::move(b->rowPtr);
    c->batch_col = std::move(b->colInd);
    wtc_batch_free(b);
    if (rc == WTC_OK && st) {
        st->prefix_ms = sum.prefix_ms;
        st->bitmap_ms = sum.bitmap_ms;
        st->count_ms = sum.count_ms;
        st->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        st->wedges = wsum;
        st->blocks = c->pipe[0].grid;
        st->launches = launches;
        st->small_graphs = small;
    }
    return rc;
}

}  // extern "C"
