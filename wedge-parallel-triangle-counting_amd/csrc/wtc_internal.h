// wtc_internal.h -- library-internal pieces shared by the C-ABI files
// (wtc_api.hip, wtc_serve.hip): device/pinned buffers, execution slots and
// the asynchronous graph-set pipeline.
#pragma once
#include <mutex>
#include <string>
#include <vector>

#include "host_util.h"
#include "wtc.h"
#include "wtc_kernels.h"
#include "wtc_launch.h"

namespace wtc {

// Records the HIP failure text for wtc_last_error_detail and returns WTC_ERR_HIP.
int hip_fail(const char *what, hipError_t e);
void set_error(const std::string &msg);

#define WTC_HIPC(call)                                            \
    do {                                                          \
        hipError_t _e = (call);                                   \
        if (_e != hipSuccess) return ::wtc::hip_fail(#call, _e);  \
    } while (0)

// Readable bytes every device buffer carries past its requested size.
constexpr size_t kDevTailSlack = 256;

// Owning, move-only buffer: freed by release() or on destruction.
template <class Free>
struct OwnedBuf {
    void *p = nullptr;
    size_t cap = 0;
    OwnedBuf() = default;
    OwnedBuf(const OwnedBuf &) = delete;
    OwnedBuf &operator=(const OwnedBuf &) = delete;
    OwnedBuf(OwnedBuf &&o) noexcept : p(o.p), cap(o.cap) {
        o.p = nullptr;
        o.cap = 0;
    }
    OwnedBuf &operator=(OwnedBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            cap = o.cap;
            o.p = nullptr;
            o.cap = 0;
        }
        return *this;
    }
    ~OwnedBuf() { release(); }
    void release() {
        if (p) Free::free(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

struct DeviceFree {
    static void free(void *p) { (void)hipFree(p); }
};
struct PinnedFree {
    static void free(void *p) { (void)hipHostFree(p); }
};

// Grow-only device buffer.
struct DevBuf : OwnedBuf<DeviceFree> {
    hipError_t ensure(size_t bytes);
};

// Grow-only pinned host buffer.
struct PinBuf : OwnedBuf<PinnedFree> {
    hipError_t ensure(size_t bytes);
};

// One asynchronous execution slot: a stream, timing events, the device
// workspace and the pinned staging of one in-flight graph-set count.
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev[5] = {};  // prefix start, bitmap start, count start, count end, split row kernel end
    hipEvent_t ev_ready = nullptr;  // input copied (waited on by the slot's stream)
    DevBuf ws, partials, tree, bm, desc, out;
    // split count: prefixes, transposition counts/totals/offsets, in-edge records
    DevBuf sp;  // split count workspace (split_layout)
    PinBuf h_desc, h_out;
    uint32_t G = 0;       // graphs of the last enqueue
    uint32_t grid = 0;    // counting grid of the last enqueue
    uint32_t rows_grid = 0;   // row-kernel grid of the last split count
    uint32_t slabs_grid = 0;  // slab-kernel grid of the last split count
    bool split = false;       // the last enqueue was a split count
    uint32_t small = 0;   // graphs of the last enqueue taken by the small-graph kernel
    int create();
    void release_device();  // frees the device workspace (keeps stream/events)
    void destroy();
};

// Counting-kernel configuration of a context.
struct Tuning {
    int num_cus = 0;
    int ways = 4;         // probes advanced together per lane
    int wpb = 4;          // waves per counting block: 1, 2, 4 or 8
    int kernel = 0;       // kKernelWalk, kKernelStaged or kKernelSplit
    int split_wpb = 8;    // waves per block of the split count's row kernel
    // the process context of the tc_GPU drop-ins takes the autotuner's
    // kernel and block shape per graph unless $WTC_COUNT_KERNEL names one
    bool auto_shape = false;
    // the split count only with a workspace the caller provides (the
    // server carves it from its HBM pool), never from the slot
    bool split_needs_base = false;
    // work split of the counting kernels: most tiles per grab (0: the
    // kernel's own default, grab_for), one stream of grabs
    uint32_t grab = 0;
    bool deal = false;    // deal super-chunks of grabs to the XCDs in turn
    uint32_t chunk = 16;  // wedges per lane staged in LDS per phase
    // In sets of two or more graphs, a graph with 2(n+1) + m <= small_words
    // is counted by the small-graph kernel from LDS (0: never);
    // small_words_max is the device's LDS per block in words (at most
    // kSmallWordsMax).
    uint32_t small_words = 0;
    uint32_t small_words_max = 0;
    // LDS per block in bytes: the split count's row bitmap (a bits) must fit
    size_t lds_bytes = 0;
    uint32_t grid_cache[2][2][4][4][kChunkMax + 2] = {};  // by kernel, offset width, ways, waves, staged
    // Most tiles per grab for a counting kernel.  The walk's default is 4:
    // on R-MAT-24 the cost per wedge varies along the order and bigger grabs
    // left the slowest waves last.  The staged kernel's is 32: it searches
    // the first tile of each grab from the tree, and with grabs of 4 that
    // search doubled its time on R-MAT-20 (7.5 vs 3.8 ms at s = 8).
    uint32_t grab_for(int k) const { return grab ? grab : k == kKernelStaged ? 32u : 4u; }
};

// Default small-graph limit of new contexts and servers, in LDS words
// (2(n+1) + m): $WTC_SMALL_GRAPH_WORDS, else kSmallWordsDefault.
constexpr uint32_t kSmallWordsDefault = 20480;
uint32_t default_small_words();
// Sets the tuning's small-graph limits for a device (want clamped to its
// LDS per block) and raises the kernel's dynamic-LDS limit to match.
int init_small_graphs(Tuning &t, const hipDeviceProp_t &prop, uint32_t want);

struct RunTimes {
    float prefix_ms = 0, bitmap_ms = 0, count_ms = 0;
    float rows_ms = 0, list_ms = 0;  // split count: its row and slab kernels
    bool split = false;
    uint64_t top_wedges = 0, in_edges = 0, list_in_edges = 0;
};

uint64_t bm_words_for(uint32_t a);
int validate_oriented_serial(const wtc_big_graph *g);
int check_args(uint32_t spread, uint32_t a);

// Host-filled descriptor of one graph (device pointers to its CSR).  ws and
// bitmap may be preset by the caller; enqueue_set assigns the missing ones
// from the slot's workspace.
GraphDesc make_desc(const void *rowPtr, const uint32_t *colInd, uint32_t n, uint32_t spread, uint32_t a,
                    uint64_t w_begin, uint64_t w_end);

// What launch_set needs to replay the launches of a staged set.
struct SetPlan {
    uint32_t G = 0, blocks = 0, max_spread = 1, chunk = 16;
    int ob = 4, ways = 4, wpb = 4, kernel = 0;
    uint32_t work = 0;  // count_work(grab, deal)
    uint64_t rows = 0, clear_words = 0;
    bool single = false;
    // small_count blocks of the small-graph kernel, whose SmallPart list
    // follows the descriptors; desc_bytes covers both
    uint32_t small_count = 0, small_lds = 0;
    uint32_t tile_graphs = 0;  // graphs counted by the wedge-parallel kernel
    size_t desc_bytes = 0;
    // split count: SplitDesc at byte split_off of the descriptor buffer
    bool split = false;
    size_t split_off = 0;
    uint32_t split_rows = 0, split_keys = 0, split_chunks = 0, split_c2cap = 0, split_c3cap = 0, split_parts = 0;
    uint32_t split_wpb = 8;
    uint64_t *split_key_part = nullptr;  // device: the key-block prefix of the transposition
    // offset width of the split kernels, and for a narrowed 8-byte CSR the
    // row offsets copied to 4 bytes at the start of the count
    int split_ob = 4;
    const uint64_t *split_rp_src = nullptr;
    uint32_t *split_rp_dst = nullptr;
    uint64_t split_rp_n = 0;
};
// Sizes the slot's workspace for the set, assigns each graph its prefix and
// bitmap (unless given), fills the pinned host descriptors and the plan.
// m and W (edges and wedges per graph, or null) let a set whose graphs all
// fit the tuning's small-graph limit take the small-graph kernel.
int stage_set(Slot &s, Tuning &t, int ob, std::vector<GraphDesc> &hd, SetPlan *plan,
              const uint64_t *m = nullptr, const uint64_t *W = nullptr, char *split_base = nullptr);
// Enqueues the set's kernels and the copy of meta + counts to the host; no
// host-to-device copies, allocations or memsets, so it can be captured into
// a graph once the descriptors are on the device.  timed: record ev[0..3].
int launch_set(Slot &s, const SetPlan &p, bool timed);
// stage_set + descriptors to the device + launch_set: the whole device
// pipeline for a graph set on the slot's stream and the copy of its counts
// into the slot's pinned buffer; returns without waiting.  All graphs of a
// set share the row-offset width ob.
int enqueue_set(Slot &s, Tuning &t, int ob, std::vector<GraphDesc> &hd, const uint64_t *m = nullptr,
                const uint64_t *W = nullptr, char *split_base = nullptr);
// The split count (a single graph with a bitmap length; wtc_split.hip):
// whether the tuning takes it for a graph, its workspace's device bytes for
// a graph of n vertices, m edges and `rows` top rows at offset width ob, and
// the SplitDesc pointers into a workspace of that size at base.  stage_set
// takes the workspace from split_base when given, else from the slot.
bool split_eligible(const Tuning &t, const GraphDesc &d, uint64_t m);
size_t split_workspace_bytes(uint32_t n, uint64_t m, uint32_t rows, int ob);
// split_layout returns the split kernels' offset width (4 for an 8-byte CSR
// whose positions fit 32 bits, split_narrow) and, through rp32, where the
// count copies its row offsets to 4 bytes (null when not narrowed).
bool split_narrow(int ob, uint64_t m);
int split_layout(SplitDesc *sd, char *base, uint32_t n, uint64_t m, uint32_t rows, int ob, uint32_t **rp32 = nullptr);
// Waits for the slot (or, if it is done, does not wait) and reads results.
int finish_set(Slot &s, uint64_t *counts, uint64_t *wedges, RunTimes *rt);
// True when the slot's last enqueue has completed.
bool slot_done(Slot &s);

// The process-wide context of the context-free entry points (tc_GPU,
// read_graph/preprocess with -p 1/2); callers hold default_ctx_mutex().
int default_context(wtc_ctx **out);
std::mutex &default_ctx_mutex();

// GPU builders behind read_graph/preprocess with -p 1/2 (wtc_prep.hip):
// the symmetric CSR of directed pairs without self-loops and ids < n, and
// the degree-ordered orientation of a symmetric CSR.  Both return host CSRs
// identical to the CPU builders'.
int gpu_symmetric_from_pairs(wtc_ctx *c, const uint32_t *pairs, uint64_t npairs, uint32_t n, wtc_big_graph **out);
int gpu_orient(wtc_ctx *c, const wtc_big_graph *sym, wtc_big_graph **out);

}  // namespace wtc

struct wtc_ctx {
    int device = 0;
    wtc::Slot slot;
    wtc::Tuning tune;
    wtc::PinBuf h_stage;              // pinned staging of uploads
    wtc::DevBuf batch_row, batch_col;  // device CSRs of wtc_batch_count_host, kept between calls
    wtc::Slot pipe[2];                 // execution slots of wtc_batch_count_host's chunks
};

struct wtc_dgraph {
    wtc_ctx *ctx;
    uint32_t n;
    uint64_t m;
    int ob;
    wtc::DevBuf rowPtr, colInd;
};
