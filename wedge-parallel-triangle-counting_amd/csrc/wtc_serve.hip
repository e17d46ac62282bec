// wtc_serve.hip -- serving: a request loop with continuous batching over
// device-resident CSRs.  No counterpart in the reference.
//
// Graphs are registered once (copied to pinned host memory); a request names
// a graph, a spread and a bitmap length.  One worker thread owns the GPU:
//   * admission: a request's device footprint (its graph's CSR if not
//     resident, plus its own wedge prefix and bitmap) is carved from a
//     fixed HBM pool; graphs stay resident after use and the least recently
//     used idle ones are evicted when the pool is short;
//   * uploads of newly needed graphs run on their own stream and overlap the
//     counting of earlier batches;
//   * whenever an execution slot is free, every admitted request whose graph
//     is resident joins one graph set and is counted by one launch sequence;
//     with several slots, a new batch starts while earlier ones are still on
//     the device;
//   * completions are detected by stream queries and handed to waiters.
// Counts are exact integers, so a request's result does not depend on the
// batch it ran in.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "wtc_internal.h"

using namespace wtc;
using Clock = std::chrono::steady_clock;

namespace {

constexpr size_t kAlign = 256;
// A batch finishes as a unit, so it only takes requests whose wedge counts
// lie within a factor kCostSpread of its oldest request's; graphs below
// kCostFloor wedges count as equal (their time is launch latency).
constexpr uint64_t kCostSpread = 16;
constexpr uint64_t kCostFloor = 1ull << 20;
size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over offsets of one device allocation, with coalescing.
class Pool {
  public:
    void reset(size_t cap) {
        free_.clear();
        if (cap) free_[0] = cap;
    }
    // Returns the offset or SIZE_MAX.
    size_t alloc(size_t bytes) {
        bytes = align_up(std::max<size_t>(bytes, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, left = it->second - bytes;
            free_.erase(it);
            if (left) free_[off + bytes] = left;
            return off;
        }
        return SIZE_MAX;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(std::max<size_t>(bytes, 1));
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::map<size_t, size_t> free_;
};

struct SrvGraph {
    uint64_t id = 0;
    uint32_t n = 0;
    uint64_t m = 0;
    PinBuf host;        // rowPtr as u32 (n+1), then colInd (m) from byte col_off
    size_t col_off = 0; // (n+1)*4 rounded up to 16 bytes: colInd 16-byte aligned
    size_t bytes = 0;   // device footprint of the CSR (with 256 bytes of slack)
    size_t off = SIZE_MAX;
    hipEvent_t uploaded = nullptr;
    bool uploading = false, resident = false;
    uint32_t pins = 0;  // admitted requests using it
    uint64_t last_use = 0;
    uint64_t first_rid = UINT64_MAX;  // oldest queued request naming it
    uint64_t W = 0;                    // wedges
    uint64_t cost = 0;                 // batching cost class: max(W, kCostFloor)
    bool split = false;                // the autotuner counts it by the split count (given a > 0)
};

enum ReqState { kQueued, kRunning, kDone, kFailed };

struct SrvReq {
    uint64_t id = 0, gid = 0;
    uint32_t spread = 0, a = 0;
    bool split = false;  // counted alone by the split count; bm_* then hold its workspace
    ReqState state = kQueued;
    size_t ws_off = SIZE_MAX, ws_bytes = 0, bm_off = SIZE_MAX, bm_bytes = 0;
    uint64_t result = 0;
    int rc = WTC_OK;
    uint32_t batch_graphs = 0;
    uint64_t batch_id = 0;
    Clock::time_point t_submit, t_start, t_done;
};

struct ExecSlot {
    Slot slot;
    bool busy = false;
    uint64_t batch_id = 0;
    std::vector<SrvReq *> reqs;
    std::vector<uint64_t> counts;
};

}  // namespace

struct wtc_server {
    int device = 0;
    wtc_server_config cfg{};
    Tuning tune;
    DevBuf pool_mem;
    Pool pool;
    hipStream_t upload = nullptr;
    std::vector<ExecSlot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::unordered_map<uint64_t, SrvGraph *> graphs;
    std::unordered_map<uint64_t, SrvReq *> reqs;
    std::deque<SrvReq *> queue;  // queued, in arrival order
    std::thread worker;
    bool stop = false;
    int fatal = WTC_OK;
    uint64_t next_gid = 1, next_rid = 1, tick = 0, batches = 0;
    wtc_server_stats stats{};
};

namespace {

char *pool_ptr(wtc_server *s, size_t off) { return s->pool_mem.as<char>() + off; }

size_t req_ws_bytes(const SrvGraph *g) { return ((size_t)g->n + 1) * sizeof(uint64_t); }
size_t req_bm_bytes(uint32_t a) { return a ? bm_words_for(a) * sizeof(uint32_t) : 0; }
// A request takes the split count when its graph's autotuned kernel is the
// split count and it asks for a bitmap; its bitmap bytes are then the split
// workspace (no bitmap in HBM).
bool req_split(const SrvGraph *g, uint32_t a) { return g->split && a > 0; }
size_t req_extra_bytes(const SrvGraph *g, uint32_t a) {
    return req_split(g, a) ? split_workspace_bytes(g->n, g->m, std::min(g->n, a), 4) : req_bm_bytes(a);
}

// Frees the CSRs of idle resident graphs, least recently used first, until
// `bytes` could be allocated for request `rid`.  Graphs that an older queued
// request still needs are kept: memory goes to requests in arrival order, so
// two requests whose graphs do not fit together cannot evict each other's
// graph back and forth.  Caller holds the lock.
size_t alloc_evicting(wtc_server *s, size_t bytes, uint64_t rid) {
    size_t off = s->pool.alloc(bytes);
    while (off == SIZE_MAX) {
        SrvGraph *victim = nullptr;
        for (auto &kv : s->graphs) {
            SrvGraph *g = kv.second;
            if (g->resident && g->pins == 0 && g->first_rid > rid &&
                (!victim || g->last_use < victim->last_use))
                victim = g;
        }
        if (!victim) return SIZE_MAX;
        s->pool.release(victim->off, victim->bytes);
        victim->off = SIZE_MAX;
        victim->resident = false;
        s->stats.evictions++;
        s->stats.resident_bytes -= victim->bytes;
        off = s->pool.alloc(bytes);
    }
    return off;
}

void finish_request(wtc_server *s, SrvReq *r, uint64_t result, int rc) {
    r->result = result;
    r->rc = rc;
    r->state = rc == WTC_OK ? kDone : kFailed;
    r->t_done = Clock::now();
    if (r->ws_off != SIZE_MAX) s->pool.release(r->ws_off, r->ws_bytes);
    if (r->bm_off != SIZE_MAX) s->pool.release(r->bm_off, r->bm_bytes);
    r->ws_off = r->bm_off = SIZE_MAX;
    auto it = s->graphs.find(r->gid);
    if (it != s->graphs.end() && it->second->pins) it->second->pins--;
    s->stats.completed++;
}

// One pass of the scheduler; returns true if anything changed.  Lock held.
bool schedule(wtc_server *s) {
    bool progress = false;
    // 1. completed batches
    for (auto &x : s->slots) {
        if (!x.busy || !slot_done(x.slot)) continue;
        x.counts.resize(x.reqs.size());
        const int rc = finish_set(x.slot, x.counts.data(), nullptr, nullptr);
        for (size_t i = 0; i < x.reqs.size(); i++) finish_request(s, x.reqs[i], x.counts[i], rc);
        x.reqs.clear();
        x.busy = false;
        progress = true;
        s->cv_done.notify_all();
    }
    // 2. uploads finished / needed
    for (auto &kv : s->graphs) {
        SrvGraph *g = kv.second;
        g->first_rid = UINT64_MAX;
        if (g->uploading && hipEventQuery(g->uploaded) == hipSuccess) {
            g->uploading = false;
            g->resident = true;
            progress = true;
        }
    }
    for (SrvReq *r : s->queue) {  // ids grow along the queue
        SrvGraph *g = s->graphs[r->gid];
        if (g->first_rid == UINT64_MAX) g->first_rid = r->id;
    }
    for (SrvReq *r : s->queue) {
        SrvGraph *g = s->graphs[r->gid];
        if (g->resident || g->uploading) continue;
        const size_t off = alloc_evicting(s, g->bytes, r->id);
        if (off == SIZE_MAX) break;  // the pool is full of graphs in use: wait for completions
        g->off = off;
        char *dst = pool_ptr(s, off);
        hipError_t e = hipMemcpyAsync(dst, g->host.p, g->col_off + g->m * 4, hipMemcpyHostToDevice, s->upload);
        if (e == hipSuccess) e = hipEventRecord(g->uploaded, s->upload);
        if (e != hipSuccess) {
            s->fatal = hip_fail("server upload", e);
            return false;
        }
        g->uploading = true;
        s->stats.uploads++;
        s->stats.resident_bytes += g->bytes;
        progress = true;
    }
    // 3. a free slot takes, oldest first, the queued requests whose graph is
    // resident and whose cost is in the class of the oldest such request
    for (auto &x : s->slots) {
        if (x.busy || s->queue.empty()) continue;
        std::vector<GraphDesc> hd;
        std::vector<uint64_t> hm, hw;  // edges and wedges per graph of hd
        uint64_t c0 = 0;
        char *split_base = nullptr;
        for (auto it = s->queue.begin(); it != s->queue.end() && hd.size() < s->cfg.max_batch;) {
            SrvReq *r = *it;
            SrvGraph *g = s->graphs[r->gid];
            if (!g->resident) {
                ++it;
                continue;
            }
            // a split request runs alone: it starts an empty batch, or waits
            if (r->split && !hd.empty()) {
                ++it;
                continue;
            }
            if (c0 == 0) c0 = g->cost;
            if (g->cost > c0 * kCostSpread || g->cost * kCostSpread < c0) {
                ++it;
                continue;
            }
            r->ws_bytes = req_ws_bytes(g);
            r->bm_bytes = req_extra_bytes(g, r->a);
            g->pins++;  // keep the graph while its own workspace is being carved out
            r->ws_off = alloc_evicting(s, r->ws_bytes, r->id);
            r->bm_off = r->bm_bytes && r->ws_off != SIZE_MAX ? alloc_evicting(s, r->bm_bytes, r->id) : SIZE_MAX;
            if (r->ws_off == SIZE_MAX || (r->bm_bytes && r->bm_off == SIZE_MAX)) {
                if (r->ws_off != SIZE_MAX) s->pool.release(r->ws_off, r->ws_bytes);
                if (r->bm_off != SIZE_MAX) s->pool.release(r->bm_off, r->bm_bytes);
                r->ws_off = r->bm_off = SIZE_MAX;
                g->pins--;
                break;  // pool exhausted: run what we have
            }
            g->last_use = ++s->tick;
            const char *base = pool_ptr(s, g->off);
            GraphDesc d = make_desc(base, reinterpret_cast<const uint32_t *>(base + g->col_off), g->n, r->spread,
                                    r->a, 0, UINT64_MAX);
            d.ws = reinterpret_cast<uint64_t *>(pool_ptr(s, r->ws_off));
            if (r->split) split_base = pool_ptr(s, r->bm_off);
            else d.bitmap = r->bm_bytes ? reinterpret_cast<uint32_t *>(pool_ptr(s, r->bm_off)) : nullptr;
            hd.push_back(d);
            hm.push_back(g->m);
            hw.push_back(g->W);
            r->state = kRunning;
            r->t_start = Clock::now();
            x.reqs.push_back(r);
            it = s->queue.erase(it);
            if (r->split) break;  // alone
        }
        if (hd.empty()) break;
        const int rc = enqueue_set(x.slot, s->tune, 4, hd, hm.data(), hw.data(), split_base);
        x.batch_id = ++s->batches;
        for (SrvReq *r : x.reqs) {
            r->batch_graphs = (uint32_t)hd.size();
            r->batch_id = x.batch_id;
        }
        s->stats.batches++;
        s->stats.max_batch_graphs = std::max<uint64_t>(s->stats.max_batch_graphs, hd.size());
        if (rc != WTC_OK) {
            for (SrvReq *r : x.reqs) finish_request(s, r, 0, rc);
            x.reqs.clear();
            s->cv_done.notify_all();
        } else {
            x.busy = true;
        }
        progress = true;
    }
    return progress;
}

void worker_main(wtc_server *s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        const bool busy = std::any_of(s->slots.begin(), s->slots.end(), [](const ExecSlot &x) { return x.busy; });
        if (s->stop && !busy && s->queue.empty()) break;
        const bool progress = s->fatal == WTC_OK && schedule(s);
        if (s->fatal != WTC_OK) {
            for (SrvReq *r : s->queue) finish_request(s, r, 0, s->fatal);
            s->queue.clear();
            s->cv_done.notify_all();
            if (s->stop) break;
        }
        if (!progress) {
            // poll the device while work is in flight; otherwise sleep until a submit
            if (busy || !s->queue.empty()) s->cv_work.wait_for(lk, std::chrono::microseconds(20));
            else s->cv_work.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
}

}  // namespace

extern "C" {

int wtc_server_create(int device, const wtc_server_config *cfg, wtc_server **out) {
    if (!out) return WTC_ERR_ARG;
    *out = nullptr;
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0) return WTC_ERR_NODEV;
    if (device < 0 || device >= cnt) return WTC_ERR_ARG;
    WTC_HIPC(hipSetDevice(device));
    wtc_server *s = new (std::nothrow) wtc_server();
    if (!s) return WTC_ERR_NOMEM;
    s->device = device;
    s->cfg.hbm_budget_bytes = cfg && cfg->hbm_budget_bytes ? cfg->hbm_budget_bytes : (8ull << 30);
    s->cfg.max_batch = cfg && cfg->max_batch ? cfg->max_batch : 256;
    s->cfg.slots = cfg && cfg->slots ? std::min<uint32_t>(cfg->slots, 8) : 2;
    s->cfg.queue_capacity = cfg && cfg->queue_capacity ? cfg->queue_capacity : 65536;
    s->cfg.probe_ways = cfg && cfg->probe_ways ? cfg->probe_ways : 4;
    s->cfg.lds_chunk = cfg && cfg->lds_chunk ? cfg->lds_chunk : 16;
    s->cfg.waves_per_block = cfg && cfg->waves_per_block ? cfg->waves_per_block : 4;
    s->cfg.small_graph_words = cfg ? cfg->small_graph_words : 0;
    if (!valid_count_waves(s->cfg.waves_per_block)) {
        delete s;
        return WTC_ERR_ARG;
    }
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete s;
        return hip_fail("hipGetDeviceProperties", e);
    }
    s->tune.num_cus = prop.multiProcessorCount;
    s->tune.ways = s->cfg.probe_ways;
    s->tune.chunk = s->cfg.lds_chunk;
    s->tune.wpb = s->cfg.waves_per_block;
    // a request whose graph the autotuner counts by the split count runs
    // alone with its workspace from the pool; sets of several graphs walk
    s->tune.kernel = kKernelSplit;
    s->tune.split_needs_base = true;
    int rc = init_small_graphs(s->tune, prop, s->cfg.small_graph_words < 0    ? 0u
                                              : s->cfg.small_graph_words == 0 ? default_small_words()
                                                                              : (uint32_t)s->cfg.small_graph_words);
    if (!rc && s->cfg.small_graph_words > 0 && (uint32_t)s->cfg.small_graph_words > s->tune.small_words_max)
        rc = WTC_ERR_ARG;
    if (rc) {
        delete s;
        return rc;
    }
    if ((e = s->pool_mem.ensure(s->cfg.hbm_budget_bytes)) != hipSuccess)
        rc = hip_fail("server pool", e);
    if (!rc && (e = hipStreamCreateWithFlags(&s->upload, hipStreamNonBlocking)) != hipSuccess)
        rc = hip_fail("server stream", e);
    s->slots.resize(s->cfg.slots);
    for (auto &x : s->slots)
        if (!rc) rc = x.slot.create();
    if (rc) {
        for (auto &x : s->slots) x.slot.destroy();
        if (s->upload) (void)hipStreamDestroy(s->upload);
        s->pool_mem.release();
        delete s;
        return rc;
    }
    s->pool.reset(s->cfg.hbm_budget_bytes);
    s->worker = std::thread(worker_main, s);
    *out = s;
    return WTC_OK;
}

int wtc_server_register(wtc_server *s, const wtc_big_graph *g, uint64_t *graph_id) {
    if (!s || !g || !graph_id || g->numVertices == 0) return WTC_ERR_ARG;
    if (g->numEdges >= 0xffffffffULL) return WTC_ERR_RANGE;
    int rc = wtc_validate_oriented(g);
    if (rc) return rc;
    SrvGraph *sg = new (std::nothrow) SrvGraph();
    if (!sg) return WTC_ERR_NOMEM;
    sg->n = g->numVertices;
    sg->m = g->numEdges;
    sg->col_off = (((size_t)sg->n + 1) * 4 + 15) & ~(size_t)15;
    sg->bytes = align_up(sg->col_off + sg->m * 4 + kDevTailSlack);
    (void)hipSetDevice(s->device);
    hipError_t e = sg->host.ensure(sg->bytes);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sg->uploaded, hipEventDisableTiming);
    if (e != hipSuccess) {
        sg->host.release();
        delete sg;
        return hip_fail("server register", e);
    }
    uint32_t *rp = sg->host.as<uint32_t>();
    uint64_t W = 0;
    for (uint64_t v = 0; v <= sg->n; v++) {
        rp[v] = (uint32_t)g->rowPtr[v];
        if (v < sg->n) W += binom2(g->rowPtr[v + 1] - g->rowPtr[v]);
    }
    sg->W = W;
    sg->cost = std::max<uint64_t>(W, kCostFloor);
    wtc_tuned tuned;
    sg->split = wtc_autotune_shape(sg->n, sg->m, W, 0, &tuned) == WTC_OK && tuned.kernel == WTC_KERNEL_SPLIT;
    if (sg->m) std::memcpy(sg->host.as<char>() + sg->col_off, g->colInd, sg->m * 4);
    std::lock_guard<std::mutex> lk(s->mu);
    sg->id = s->next_gid++;
    s->graphs[sg->id] = sg;
    s->stats.graphs++;
    *graph_id = sg->id;
    return WTC_OK;
}

int wtc_server_submit(wtc_server *s, uint64_t graph_id, uint32_t spread, uint32_t a, uint64_t *request_id) {
    if (!s || !request_id) return WTC_ERR_ARG;
    int rc = check_args(spread, a);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop || s->fatal) return WTC_ERR_BUSY;
    auto it = s->graphs.find(graph_id);
    if (it == s->graphs.end()) return WTC_ERR_ARG;
    const SrvGraph *g = it->second;
    if (align_up(g->bytes) + align_up(req_ws_bytes(g)) + align_up(req_extra_bytes(g, a)) > s->cfg.hbm_budget_bytes)
        return WTC_ERR_ADMIT;
    if (s->queue.size() >= s->cfg.queue_capacity) {
        s->stats.rejected++;
        return WTC_ERR_BUSY;
    }
    SrvReq *r = new (std::nothrow) SrvReq();
    if (!r) return WTC_ERR_NOMEM;
    r->id = s->next_rid++;
    r->gid = graph_id;
    r->spread = spread;
    r->a = a;
    r->split = req_split(g, a);
    r->t_submit = Clock::now();
    s->reqs[r->id] = r;
    s->queue.push_back(r);
    s->stats.submitted++;
    *request_id = r->id;
    s->cv_work.notify_one();
    return WTC_OK;
}

int wtc_server_wait(wtc_server *s, uint64_t request_id, int64_t timeout_us, uint64_t *triangles,
                    wtc_request_stats *st) {
    if (!s) return WTC_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request_id);
    if (it == s->reqs.end()) return WTC_ERR_ARG;
    SrvReq *r = it->second;
    auto done = [&] { return r->state == kDone || r->state == kFailed; };
    if (timeout_us < 0) s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), done)) return WTC_ERR_BUSY;
    const int rc = r->rc;
    if (triangles) *triangles = r->result;
    if (st) {
        using ms = std::chrono::duration<double, std::milli>;
        st->queue_ms = ms(r->t_start - r->t_submit).count();
        st->service_ms = ms(r->t_done - r->t_start).count();
        st->latency_ms = ms(r->t_done - r->t_submit).count();
        st->batch_graphs = r->batch_graphs;
        st->batch_id = r->batch_id;
    }
    s->reqs.erase(it);
    delete r;
    return rc;
}

int wtc_server_get_stats(wtc_server *s, wtc_server_stats *st) {
    if (!s || !st) return WTC_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->stats;
    st->queued = s->queue.size();
    uint64_t res = 0;
    for (auto &kv : s->graphs) res += kv.second->resident ? 1 : 0;
    st->resident_graphs = res;
    return s->fatal;
}

void wtc_server_destroy(wtc_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device);
    for (auto &x : s->slots) x.slot.destroy();
    if (s->upload) {
        (void)hipStreamSynchronize(s->upload);
        (void)hipStreamDestroy(s->upload);
    }
    for (auto &kv : s->graphs) {
        if (kv.second->uploaded) (void)hipEventDestroy(kv.second->uploaded);
        kv.second->host.release();
        delete kv.second;
    }
    for (auto &kv : s->reqs) delete kv.second;
    s->pool_mem.release();
    delete s;
}

}  // extern "C"
