// wtc_kernels.h -- device data layout shared by the HIP kernels and the host
// code that launches them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wtc {

constexpr int kWave = 64;               // CDNA wavefront
constexpr int kWavesPerBlock = 4;       // counting-kernel block = 256 lanes
constexpr int kBlock = kWave * kWavesPerBlock;
constexpr int kCountWavesMax = 8;       // counting-kernel block: 1..8 waves, chosen per launch
constexpr uint32_t kChunkMax = 16;      // wedges per lane staged in LDS at once
constexpr int kTreeLevels = 12;         // cached levels of the wedge-prefix search tree
constexpr uint32_t kTreeSize = (1u << kTreeLevels) - 1u;
constexpr int kScanItems = 16;          // vertices per thread in the prefix scan
constexpr uint32_t kScanChunk = kBlock * kScanItems;  // vertices per scan block
constexpr uint32_t kSentinel = 0xffffffffu;           // empty LDS wedge slot
constexpr int kSmallBlock = 512;        // small-graph kernel block: one graph per block
// Largest LDS footprint, in 32-bit words, of a graph the small-graph kernel
// takes: 2(n+1) + m <= 2^16 keeps m < 2^16, so W <= C(m,2) < 2^31 and every
// wedge index of such a graph fits 32 bits.
constexpr uint32_t kSmallWordsMax = 1u << 16;

// Wedges per block the host aims at when it splits a graph of a small-graph
// set over several blocks of the small-graph kernel.
constexpr uint64_t kSmallBlockWedges = 16384;

// One block of the small-graph kernel: wedges [wcount*part/parts,
// wcount*(part+1)/parts) of graph g's range.
struct SmallPart {
    uint32_t g;
    uint16_t part, parts;
};
static_assert(sizeof(SmallPart) == 8, "SmallPart layout");

// GraphDesc::mode
enum : uint32_t {
    kModeTiles = 0,  // counted by k_count in tiles of 64*s wedges
    kModeSmall = 1,  // counted by k_count_small, one block with the CSR in LDS
    kModeInvalid = 2,  // failed the device-side CSR check: nothing is read
                       // beyond rowPtr, and its count is set to kCountInvalid
};

// Count of a graph whose CSR failed the device-side check.
constexpr unsigned long long kCountInvalid = ~0ull;

// One graph of a graph set.  A graph set is any number of oriented CSRs
// resident in HBM (ids local to each graph, row offsets relative to the
// graph's own colInd); one count launch covers all of them.  A single graph
// is a set of one.  The host fills the first block of fields, the setup
// kernel the last three.
struct GraphDesc {
    // host-filled
    const void *rowPtr;      // n+1 offsets of the set's offset width
    const uint32_t *colInd;  // ids in [0, n), every row strictly ascending
    uint64_t *ws;            // n+1 wedge prefix (workspace slice)
    uint32_t *bitmap;        // ceil(C(a,2)/32) words (workspace slice), or null
    uint64_t bm_words;       // a*a/32: words of the graph's top-a bitmap
    uint64_t row_base;       // first bitmap row of this graph in the set's row list
    uint64_t blk_base;       // first prefix-scan block of this graph
    uint64_t bm_clear_base;  // first bitmap word of this graph in the set's clear list
    uint64_t w_begin;        // wedge range to count, relative to the graph
    uint64_t w_end;
    uint32_t n;              // vertices
    uint32_t spread;         // wedges per lane (s)
    uint32_t a;              // bitmap length (top-a ranks), multiple of 32
    uint32_t thr;            // bitmap threshold max(n,a) - a (a = 0: n)
    uint32_t bm_rows;        // min(n, a)
    uint32_t mode;           // kModeTiles, kModeSmall or kModeInvalid
    // device-filled by the setup kernel
    uint64_t wbase;          // first counted wedge (graph-local index)
    uint64_t wcount;         // wedges counted
    uint64_t tile_base;      // first tile of this graph
};
static_assert(sizeof(GraphDesc) == 128, "GraphDesc layout");

// meta[] slots written by the setup kernel
// (kMetaNext..kMetaNext+7: per-XCD work counters of the counting kernel).
// The split count: kMetaRows and kMetaSlabs, tile counters of its row and
// slab kernels; kMetaTopWedges and kMetaListWedges, the wedges each of them
// checked; kMetaInEdges, the top rows' in-edges; kMetaAllEdges, every
// grouped in-edge.
enum : int {
    kMetaTiles = 0,
    kMetaWedges = 1,
    kMetaRows = 2,
    kMetaTopWedges = 3,
    kMetaInEdges = 4,
    kMetaSlabs = 5,
    kMetaListWedges = 6,
    kMetaAllEdges = 7,
    kMetaNext = 8,
    kMetaSlots = 16
};

// Split count of one graph (wtc_split.hip).  The in-edges are grouped by key
// in three counting-sort levels: by part of kSplitPart keys, by sub-part of
// kSplitSub keys (kSplitFan2 per part), by key.  Each level's scatter keeps
// at most kSplitFan2 or kSplitSub output runs open per block, so that the
// lines of a run fill in the L2 cache before they are written back.
// kSplitChunk2/3: records per block of the second and third levels;
// kSplitBlock: threads per transposition block; kSplitTile: in-edges per
// tile of the row kernel.
constexpr uint32_t kSplitPart = 32768;
constexpr uint32_t kSplitPartsMax = 256;
constexpr uint32_t kSplitSub = 256;
constexpr uint32_t kSplitSubShift = 8;
constexpr uint32_t kSplitFan2 = kSplitPart / kSplitSub;
static_assert(kSplitSub == 1u << kSplitSubShift && kSplitFan2 * kSplitSub == kSplitPart, "key levels");
static_assert(kSplitPart <= 65536 && kSplitSub <= 256, "A keys fit 16 bits, B keys 8");
constexpr uint32_t kSplitChunk2 = 65536;
constexpr uint32_t kSplitChunk3 = 65536;
constexpr int kSplitBlock = 1024;
constexpr uint32_t kSplitTile = 1024;
// Keys per block of the key-offset scan (four per thread).
constexpr uint32_t kKeyChunk = 4 * kSplitBlock;
// Slabs of the list rows (middle vertex below thr): slab s holds the rows
// whose adjacency starts in colInd[s*kSlabSpan, (s+1)*kSlabSpan).  The slab
// kernel hashes colInd[s*kSlabSpan, +kSlabWin) into LDS, one bucket of
// kSlabWays slots per position, which covers every such row of out-degree
// <= kSlabRowMax (longer rows are searched in HBM).
constexpr uint32_t kSlabSpan = 2048;
constexpr uint32_t kSlabRowMax = 1024;
constexpr uint32_t kSlabWin = kSlabSpan + kSlabRowMax;
constexpr uint32_t kSlabWays = 4;
constexpr uint32_t kSlabTile = 1024;

// Workspace and device-filled state of a split count.  The top rows are the
// ranks [thr, n) of the graph's GraphDesc; OffT arrays use the set's offset
// width.  Every in-edge (v, j) (j a position of adj+(v) other than its last)
// is grouped by a key: w - thr for w = colInd[j] >= thr (a top row), else
// rows + the slab of w.  Records (key, pair, slab_aux) go through the
// levels A -> B -> ie: A bucketed by part, B by sub-part, ie by key (the
// pairs only, in A's pair array; a list record's slab_aux in A's aux
// array).  A record's position names its part (in A) or sub-part (in B), so
// A keeps the key's low 15 bits and B its low 8.  A top record's pair is the tail [lo, hi) of colInd; a list
// record's is (j, hi), its tail [j + 1, hi) (the one in-edge whose tail the
// range cuts at the front: [clip_lo, hi)).
struct SplitDesc {
    uint64_t *ws_full;    // n+1: the full wedge prefix
    uint64_t *ie_pre;     // n+1: prefix of the in-edges (positions of adj+(v) but the last) of each row
    uint64_t *hist1;      // chunks x parts: first-level counts, then record offsets
    uint64_t *part_base;  // parts+1: first record of each part; part_base[parts] = the in-edges
    uint32_t *c2base;     // parts+1: first second-level block of each part
    uint32_t *hist2;      // c2cap x kSplitFan2: second-level counts, then column prefixes
    uint64_t *sub_base;   // subs+1: first record of each sub-part
    uint32_t *c3base;     // subs+1: first third-level block of each sub-part
    uint32_t *hist3;      // c3cap x kSplitSub: third-level counts, then column prefixes
    uint32_t *sub_tot;    // subs: records of each sub-part
    uint32_t *list_bid;   // parts: the first list vertex w whose key lies in each part (0 up to the first list part)
    uint32_t *slab_row;   // slabs+1: the first vertex w with rowPtr[w] >= s*kSlabSpan (set for s*kSlabSpan <= rowPtr[thr])
    uint64_t *part_tot;   // parts: records of each part
    uint64_t *key_part;   // key blocks: in-edges of each block of kKeyChunk keys, then their prefix
    uint32_t *key_tot;    // keys: in-edges of each key
    uint16_t *rec_id;     // cap: A keys within their part (key mod kSplitPart)
    void *rec;            // cap pairs of OffT: A pairs, then ie
    uint32_t *rec_aux;    // cap: A slab_aux words (list records), then ie_aux
    uint8_t *rec_id2;     // cap: B keys within their sub-part (key mod kSplitSub)
    void *rec2;           // cap pairs of OffT: B pairs, then the packed top ids (three 21-bit ids per word)
    uint32_t *rec_aux2;   // cap: B slab_aux words
    void *rp_in;          // keys+1 OffT: first in-edge of each key
    void *ie;             // = rec: the records' pairs grouped by key
    uint32_t *ie_aux;     // = rec_aux: the list records' slab_aux words, beside ie
    uint64_t cap;         // records each array holds (>= the in-edges)
    uint32_t rows;        // top rows: n - thr
    uint32_t keys;        // rows + slabs
    uint32_t chunks;      // vertex chunks of the first level (about equal shares of the in-edges)
    uint32_t parts;       // parts of kSplitPart keys (<= kSplitPartsMax)
    uint32_t subs;        // sub-parts: parts * kSplitFan2
    uint32_t slabs;       // slabs of the list rows: ceil(m / kSlabSpan)
    uint32_t c2cap, c3cap;  // second- and third-level blocks the hists hold (>= those the records need)
    // device-filled by the range kernel
    uint64_t w_lo, w_hi;  // the counted wedges of the full order, clamped to [0, W]
    uint64_t clip_j;      // the in-edge whose tail the range cuts at the front (~0: none)
    uint64_t clip_lo;     // that tail's first position
    uint32_t v_lo, v_hi;  // vertices holding w_lo and w_hi - 1 (v_lo > v_hi: none)
    uint32_t clip;        // 1 unless [w_lo, w_hi) is the whole order
    uint32_t pad_;
};
static_assert(sizeof(SplitDesc) == 280, "SplitDesc layout");

// Where adj+(w) of a list row lies in its slab's window: bits 0-12 its start
// rowPtr[w] - slab*kSlabSpan, bits 13-25 its length (kSlabAuxLong: longer
// than kSlabRowMax, searched in HBM).
constexpr uint32_t kSlabAuxLong = 0x1fffu;
static_assert(kSlabSpan <= 8192 && kSlabRowMax < kSlabAuxLong, "slab_aux fields");
static_assert(sizeof(SplitDesc) % 8 == 0, "SplitDesc alignment");

}  // namespace wtc
