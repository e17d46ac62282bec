// wtc_launch.h -- host-callable launchers for the kernels in wtc_kernels.hip.
// `ob` is the byte width of the row offsets of the graph set (4 or 8).
#pragma once
#include "wtc_kernels.h"

namespace wtc {

hipError_t launch_debug_decode(const uint32_t *d, const uint64_t *iv, uint64_t k, uint32_t *w, uint32_t *u,
                               hipStream_t st);
hipError_t launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t st);
hipError_t launch_fill(uint32_t *p, uint64_t words, uint32_t v, hipStream_t st);

// Exclusive scan of partials[0..nb) in place (one block).
hipError_t launch_scan_partials(uint64_t *partials, uint32_t nb, hipStream_t st);
// Prefix-scan blocks one graph of n vertices needs.
uint32_t prefix_blocks(uint32_t n);
// Wedge prefix of every graph of the set: nblocks = sum of prefix_blocks(n),
// partials holds nblocks entries.
hipError_t launch_prefix(const GraphDesc *desc, uint32_t G, int ob, uint32_t nblocks, uint64_t *partials,
                         hipStream_t st);
hipError_t launch_setup(GraphDesc *desc, uint32_t G, uint64_t *tree, uint64_t *meta, unsigned long long *counts,
                        hipStream_t st);
// Clears every graph's bitmap (clear_words = sum of their word counts, laid
// out by bm_clear_base) and sets its bits; rows = sum of bm_rows.
hipError_t launch_bitmap(const GraphDesc *desc, uint32_t G, int ob, uint64_t rows, uint64_t clear_words,
                         hipStream_t st);
// Counts the mode-small graphs named by parts[0..nparts) with one block per
// part; lds_bytes (<= 4*kSmallWordsMax) must hold 4*(2(n+1) + m) for each
// of them.
hipError_t launch_count_small(const GraphDesc *desc, const SmallPart *parts, uint32_t nparts, int ob,
                              unsigned long long *counts, uint32_t lds_bytes, hipStream_t st);
// Raises the small-graph kernel's dynamic-LDS limit to lds_bytes.
hipError_t allow_small_lds(uint32_t lds_bytes);
// LDS bytes of a counting block of wpb waves that stages `staged` wedges
// per lane.
size_t count_lds_bytes(uint32_t staged, int wpb);
// Waves per counting block the kernel accepts: 1, 2, 4 or 8.
bool valid_count_waves(int wpb);
// Counting kernels: kKernelWalk (k_walk: waves walk the wedge order in steps
// of 64 consecutive wedges) and kKernelStaged (k_count: per-lane spreads
// staged through LDS).  Both count the same wedges of the same tiles.
// kKernelSplit: a single graph with a > 0 counted by the split count (its
// in-edges grouped by middle vertex: the top-a rows against one LDS bitmap
// row, the rest against LDS windows of colInd; wtc_split.hip); any other set
// with it by k_walk.
enum : int { kKernelWalk = 0, kKernelStaged = 1, kKernelSplit = 2 };
// `ways`: independent probes each lane advances together (1, 2, 4 or 8; for
// k_walk, the steps in flight); `wpb`: waves per block; `staged` = min(max
// spread, chunk) (k_count's LDS staging).
hipError_t count_grid(int kernel, int ob, uint32_t staged, int ways, int wpb, int num_cus, uint32_t *grid);
// `chunk` (1..kChunkMax): wedges per lane staged in LDS per phase (k_count).
// `work`: bits 0-15 the most tiles per grab (0: 32), bit 16 set to deal
// super-chunks of grabs to the 8 XCDs in turn (else one stream of grabs).
hipError_t launch_count(int kernel, const GraphDesc *desc, uint32_t G, int ob, const uint64_t *tree, uint64_t *meta,
                        unsigned long long *counts, uint32_t max_spread, uint32_t chunk, int ways, int wpb,
                        uint32_t work, uint32_t grid, hipStream_t st);
// Split count (wtc_split.hip).  partials holds 2*nblocks entries.
hipError_t launch_split_prefix(const GraphDesc *desc, const SplitDesc *sd, int ob, uint32_t nblocks,
                               uint64_t *partials, hipStream_t st);
// The counted range's ends and the in-edge it cuts at the front.
hipError_t launch_split_range(const GraphDesc *desc, SplitDesc *sd, int ob, hipStream_t st);
// dst[i] = src[i] for i < count (8-byte row offsets below 2^32 to 4 bytes).
hipError_t launch_split_narrow_offsets(const uint64_t *src, uint32_t *dst, uint64_t count, hipStream_t st);
// The three-level counting sort of the in-edges by key (chunks first-level
// blocks; c2cap and c3cap bound the second- and third-level blocks; parts
// parts, keys keys; key_part = the SplitDesc's key_part, on the device);
// zeroes the tile counters, the wedge tallies and counts[0].
hipError_t launch_split_transpose(const GraphDesc *desc, const SplitDesc *sd, int ob, uint32_t chunks,
                                  uint32_t c2cap, uint32_t c3cap, uint32_t parts, uint32_t keys, uint64_t *key_part,
                                  uint64_t *meta, unsigned long long *counts, hipStream_t st);
// Waves per block of the row kernel: 4, 8 or 16.
bool valid_split_waves(int wpb);
// LDS bytes of the row kernel's bitmap for `rows` top rows.
size_t split_rows_lds(uint32_t rows);
// Persistent grid of the row kernel for `rows` top rows (its LDS bitmap).
hipError_t split_rows_grid(int ob, uint32_t rows, int wpb, int num_cus, uint32_t *grid);
hipError_t launch_split_rows(const GraphDesc *desc, const SplitDesc *sd, int ob, uint32_t rows, int wpb,
                             uint64_t *meta, unsigned long long *counts, uint32_t grid, hipStream_t st);
// Persistent grid of the slab kernel (the list rows).
hipError_t split_slabs_grid(int ob, int num_cus, uint32_t *grid);
hipError_t launch_split_slabs(const GraphDesc *desc, const SplitDesc *sd, int ob, uint64_t *meta,
                              unsigned long long *counts, uint32_t grid, hipStream_t st);
// Vertex chunks of the transposition (equal shares of the in-edges): one
// per kSplitChunkVerts vertices, at least 1 and at most kSplitChunksMax.
constexpr uint32_t kSplitChunksMax = 1024;
static_assert(kSplitChunksMax <= (uint32_t)kSplitBlock, "one first-level chunk per thread of the part scan");
constexpr uint32_t kSplitChunkVerts = 2048;

// Fewest wedges a grab of the counting kernels covers: every grab is one
// atomic on a shared counter, so a grab of a few 64-wedge tiles (spread 1)
// makes the atomics the bound (R-MAT-20 at s = 1 with 4 tiles per grab:
// 57.7 ms against 25.8 with 32).
constexpr uint32_t kGrabWedgesMin = 2048;

// The `work` word of a grab size and XCD dealing; a nonzero grab size is
// raised to kGrabWedgesMin wedges of tiles of 64 * spread wedges (at most
// 32 tiles).
inline uint32_t count_work(uint32_t tiles_per_grab, bool deal, uint32_t spread = 1) {
    if (tiles_per_grab) {
        const uint32_t per_tile = 64u * (spread ? spread : 1u);
        const uint32_t need = (kGrabWedgesMin + per_tile - 1) / per_tile;
        if (tiles_per_grab < need) tiles_per_grab = need < 32u ? need : 32u;
    }
    return (tiles_per_grab & 0xffffu) | (deal ? 1u << 16 : 0u);
}

}  // namespace wtc
