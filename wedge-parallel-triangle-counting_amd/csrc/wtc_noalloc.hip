// wtc_noalloc.hip -- the count of a caller-owned device CSR without any
// device allocation: every workspace buffer of the count (wedge prefix, scan
// partials, search tree, top-a bitmap, descriptor, work counters) lives in
// one fixed per-device array, `scratch`, that the library holds for the life
// of the process.  For callers that cannot allocate: code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc.
//
// The entry point enqueues kernels only (no hipMalloc/hipFree, no memset, no
// host-to-device copy): the descriptor is written by a kernel that takes it
// by value, everything that must start at zero is cleared by the kernels
// themselves, and the count is accumulated in the caller's output word.
#include <mutex>

#include "wtc_internal.h"

namespace wtc {

// Largest workspace of one call: 8(n+1) prefix + 8 per 4096 vertices of scan
// partials + a*a/8 bytes of bitmap + the tree, descriptor and counters (each
// rounded to 256 B) must fit.  R-MAT-24 (n = 2^24) fits with a up to 16384.
constexpr size_t kNoAllocScratchBytes = 256ull << 20;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

// Device address of `scratch` on each device (the host-side symbol is not a
// device address) and that device's counting-kernel grid per offset width.
struct NoAllocDevice {
    unsigned char *ptr = nullptr;
    int num_cus = 0;
    uint32_t grid[2][kChunkMax + 1] = {};  // by offset width and wedges staged per lane
};
std::mutex g_mu;
NoAllocDevice g_dev[kMaxDevices];

constexpr int kNoAllocWays = 4;
constexpr uint32_t kNoAllocChunk = 12;
constexpr int kNoAllocWaves = 4;

// Device-side CSR check of a caller-owned graph (the device form of
// wtc_validate_oriented): rowPtr[0] = 0, rowPtr[n] = m, offsets monotone,
// every row strictly ascending with ids above the row and below n.  A row's
// ids are read only after its two offsets are known to lie in [0, m].  On any
// failure the descriptor's mode becomes kModeInvalid, which the later kernels
// of the launch honour (no tiles, no bitmap rows, count kCountInvalid).
template <typename OffT>
__global__ __launch_bounds__(kBlock) void k_check_csr(GraphDesc *__restrict__ desc, uint64_t m) {
    GraphDesc &D = desc[0];
    const OffT *__restrict__ rowPtr = static_cast<const OffT *>(D.rowPtr);
    const uint32_t *__restrict__ colInd = D.colInd;
    const uint32_t n = D.n;
    bool bad = false;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < n && !bad; v += stride) {
        const uint64_t s = rowPtr[v], e = rowPtr[v + 1];
        if ((v == 0 && s != 0) || (v + 1 == n && e != m) || e < s || e > m) {
            bad = true;
            break;
        }
        uint64_t prev = v;
        for (uint64_t j = s; j < e; j++) {
            const uint64_t x = colInd[j];
            if (x <= prev || x >= n) {
                bad = true;
                break;
            }
            prev = x;
        }
    }
    if (bad) D.mode = kModeInvalid;  // every failing lane stores the same value
}

__global__ void k_store_desc(GraphDesc d, GraphDesc *__restrict__ dst) {
    if (threadIdx.x == 0) dst[0] = d;
}

size_t round256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

}  // namespace wtc

using namespace wtc;

extern "C" int wtc_count_noalloc(const void *d_row_ptr, uint32_t offset_bytes, const uint32_t *d_col_ind,
                                 uint32_t n, uint64_t m, uint32_t spread, uint32_t adjacency_matrix_len,
                                 uint64_t w_begin, uint64_t w_end, uint64_t *d_triangles, void *stream) {
    const uint32_t a = adjacency_matrix_len;
    int rc = check_args(spread, a);
    if (rc) return rc;
    if (!d_row_ptr || (m && !d_col_ind) || !d_triangles || n == 0) return WTC_ERR_ARG;
    if (offset_bytes != 4 && offset_bytes != 8) return WTC_ERR_ARG;
    if (offset_bytes == 4 && m >= 0xffffffffULL) return WTC_ERR_RANGE;

    // workspace layout inside scratch, each piece on a 256-byte boundary
    const uint32_t nb = prefix_blocks(n);
    const size_t off_meta = 0;
    const size_t off_desc = off_meta + round256((size_t)kMetaSlots * sizeof(uint64_t));
    const size_t off_tree = off_desc + round256(sizeof(GraphDesc));
    const size_t off_part = off_tree + round256((size_t)kTreeSize * sizeof(uint64_t));
    const size_t off_ws = off_part + round256((size_t)nb * sizeof(uint64_t));
    const size_t off_bm = off_ws + round256(((size_t)n + 1) * sizeof(uint64_t));
    const uint64_t bm_words = a ? bm_words_for(a) : 0;
    const size_t need = off_bm + round256((size_t)bm_words * sizeof(uint32_t));
    if (need > kNoAllocScratchBytes) {
        set_error("wtc_count_noalloc: workspace of " + std::to_string(need) + " bytes exceeds the " +
                  std::to_string(kNoAllocScratchBytes) + "-byte scratch array");
        return WTC_ERR_RANGE;
    }

    int dev = 0;
    WTC_HIPC(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return WTC_ERR_ARG;
    NoAllocDevice nd;
    {
        // once per device, with that device current: the device address of
        // this device's copy of the array, and the counting grid
        std::lock_guard<std::mutex> lk(g_mu);
        NoAllocDevice &x = g_dev[dev];
        if (!x.ptr) {
            void *p = nullptr;
            WTC_HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            if (!p) return hip_fail("hipGetSymbolAddress(scratch)", hipErrorInvalidSymbol);
            int cus = 0;
            WTC_HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            x.num_cus = cus;
            x.ptr = static_cast<unsigned char *>(p);
        }
        const int oi = offset_bytes == 8 ? 1 : 0;
        const uint32_t staged = std::min(spread, kNoAllocChunk);
        if (!x.grid[oi][staged])
            WTC_HIPC(count_grid(kKernelWalk, (int)offset_bytes, staged, kNoAllocWays, kNoAllocWaves, x.num_cus,
                                &x.grid[oi][staged]));
        nd = x;
    }
    const uint32_t grid = nd.grid[offset_bytes == 8 ? 1 : 0][std::min(spread, kNoAllocChunk)];

    unsigned char *base = nd.ptr;
    uint64_t *meta = reinterpret_cast<uint64_t *>(base + off_meta);
    GraphDesc *ddesc = reinterpret_cast<GraphDesc *>(base + off_desc);
    uint64_t *tree = reinterpret_cast<uint64_t *>(base + off_tree);
    uint64_t *partials = reinterpret_cast<uint64_t *>(base + off_part);
    GraphDesc d = make_desc(d_row_ptr, d_col_ind, n, spread, a, w_begin, w_end);
    d.mode = kModeTiles;
    d.ws = reinterpret_cast<uint64_t *>(base + off_ws);
    d.bitmap = a ? reinterpret_cast<uint32_t *>(base + off_bm) : nullptr;
    d.bm_clear_base = 0;
    d.row_base = 0;
    d.blk_base = 0;
    unsigned long long *counts = reinterpret_cast<unsigned long long *>(d_triangles);
    hipStream_t st = static_cast<hipStream_t>(stream);

    k_store_desc<<<dim3(1), dim3(64), 0, st>>>(d, ddesc);
    WTC_HIPC(hipGetLastError());
    const uint64_t cb = ((uint64_t)n + kBlock - 1) / kBlock;
    const dim3 cgrid((uint32_t)(cb < 8192 ? cb : 8192));
    if (offset_bytes == 8) k_check_csr<uint64_t><<<cgrid, dim3(kBlock), 0, st>>>(ddesc, m);
    else k_check_csr<uint32_t><<<cgrid, dim3(kBlock), 0, st>>>(ddesc, m);
    WTC_HIPC(hipGetLastError());
    WTC_HIPC(launch_prefix(ddesc, 1, (int)offset_bytes, nb, partials, st));
    WTC_HIPC(launch_setup(ddesc, 1, tree, meta, counts, st));
    WTC_HIPC(launch_bitmap(ddesc, 1, (int)offset_bytes, d.bm_rows, bm_words, st));
    WTC_HIPC(launch_count(kKernelWalk, ddesc, 1, (int)offset_bytes, tree, meta, counts, spread, kNoAllocChunk, kNoAllocWays, kNoAllocWaves,
                          count_work(0, true),
                          grid, st));
    return WTC_OK;
}
