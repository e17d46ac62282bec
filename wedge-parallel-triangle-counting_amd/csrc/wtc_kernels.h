// wtc_kernels.h -- device data layout shared by the HIP kernels and the host
// code that launches them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wtc {

constexpr int kWave = 64;               // CDNA wavefront
constexpr int kWavesPerBlock = 4;       // counting-kernel block = 256 lanes
constexpr int kBlock = kWave * kWavesPerBlock;
constexpr uint32_t kChunkMax = 16;      // wedges per lane staged in LDS at once
constexpr int kTreeLevels = 12;         // cached levels of the wedge-prefix search tree
constexpr uint32_t kTreeSize = (1u << kTreeLevels) - 1u;
constexpr int kScanItems = 16;          // vertices per thread in the prefix scan
constexpr uint32_t kScanChunk = kBlock * kScanItems;  // vertices per scan block
constexpr uint32_t kSentinel = 0xffffffffu;           // empty LDS wedge slot

// One graph of a graph set.  A graph set is a disjoint union of oriented
// CSRs laid out back to back in HBM (vertex ids shifted by vbase, edge
// offsets by the rows before); a single graph is a set of one.  The host
// fills the first block of fields, the setup kernel the rest.
struct GraphDesc {
    // host-filled
    uint64_t bm_size;      // C(a,2): bits of the graph's top-a bitmap
    uint64_t bm_word_off;  // first 32-bit word of that bitmap in the pool
    uint64_t row_base;     // first bitmap row of this graph in the set's row list
    uint64_t w_begin;      // wedge range to count, relative to the graph
    uint64_t w_end;
    uint32_t vbase;        // first vertex id of the graph in the union
    uint32_t n;            // vertices
    uint32_t spread;       // wedges per lane (s)
    uint32_t a;            // bitmap length (top-a ranks), multiple of 32
    uint32_t thr;          // union id threshold for the bitmap: vbase + max(n,a) - a
    uint32_t bm_rows;      // min(n, a)
    // device-filled by the setup kernel
    uint64_t wbase;        // union wedge index of the first counted wedge
    uint64_t wcount;       // wedges counted
    uint64_t tile_base;    // first tile of this graph
};
static_assert(sizeof(GraphDesc) == 88, "GraphDesc layout");

// meta[] slots written by the setup kernel
enum : int { kMetaTiles = 0, kMetaWedges = 1, kMetaSlots = 4 };

}  // namespace wtc
