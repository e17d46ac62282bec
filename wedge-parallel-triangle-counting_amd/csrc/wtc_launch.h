// wtc_launch.h -- host-callable launchers for the kernels in wtc_kernels.hip.
// `ob` is the byte width of the row offsets (4 or 8).
#pragma once
#include "wtc_kernels.h"

namespace wtc {

hipError_t launch_debug_decode(const uint32_t *d, const uint64_t *iv, uint64_t k, uint32_t *w, uint32_t *u,
                               hipStream_t st);
uint32_t prefix_blocks(uint32_t N);
hipError_t launch_prefix(const void *rowPtr, int ob, uint32_t N, uint64_t *partials, uint64_t *ws, hipStream_t st);
hipError_t launch_setup(const uint64_t *ws, GraphDesc *desc, uint32_t G, uint64_t *tree, uint64_t *meta,
                        unsigned long long *counts, hipStream_t st);
hipError_t launch_bitmap(const void *rowPtr, int ob, const uint32_t *colInd, const GraphDesc *desc, uint32_t G,
                         uint64_t rows, uint32_t *bm, uint64_t words, hipStream_t st);
size_t count_lds_bytes(uint32_t max_spread);
hipError_t count_grid(int ob, uint32_t max_spread, int num_cus, uint32_t *grid);
hipError_t launch_count(const void *rowPtr, int ob, const uint32_t *colInd, const uint64_t *ws, const uint64_t *tree,
                        const GraphDesc *desc, uint32_t G, const uint64_t *meta, const uint32_t *bm,
                        unsigned long long *counts, uint32_t max_spread, uint32_t grid, hipStream_t st);

}  // namespace wtc
