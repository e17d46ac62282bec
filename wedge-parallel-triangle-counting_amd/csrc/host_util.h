// host_util.h -- small host-side helpers shared by the library's C++ files:
// a deterministic parallel-for over index ranges, a counter-based RNG and
// allocation helpers that report failure instead of throwing.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <new>
#include <thread>
#include <vector>

#include "wtc.h"

namespace wtc {

// Worker threads for host preprocessing: WTC_HOST_THREADS, else
// OMP_NUM_THREADS, else the online cores (capped at 64).
int host_threads();

// Runs f(begin, end, thread_index) on [0, n) split into `chunks` contiguous
// pieces claimed dynamically by host_threads() workers.  The split points do
// not depend on the thread count, so results written per chunk are
// deterministic.
template <class F>
void par_chunks(uint64_t n, uint64_t chunk, F &&f) {
    if (n == 0) return;
    if (chunk == 0) chunk = 1;
    uint64_t nchunks = (n + chunk - 1) / chunk;
    int nt = (int)std::min<uint64_t>((uint64_t)host_threads(), nchunks);
    std::atomic<uint64_t> next{0};
    auto body = [&](int tid) {
        for (;;) {
            uint64_t c = next.fetch_add(1, std::memory_order_relaxed);
            if (c >= nchunks) break;
            uint64_t b = c * chunk, e = std::min(n, b + chunk);
            f(b, e, tid);
        }
    };
    if (nt <= 1) {
        body(0);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (int t = 1; t < nt; t++) th.emplace_back(body, t);
    body(0);
    for (auto &t : th) t.join();
}

// SplitMix64 finaliser; rng(seed, stream, i) is a counter-based generator:
// the i-th draw of a stream depends only on (seed, stream, i).
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static inline uint64_t rng(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix64(mix64(seed * 0x9e3779b97f4a7c15ULL + stream) + i * 0x9e3779b97f4a7c15ULL);
}
static inline double rng_unit(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

template <class T>
T *alloc_n(uint64_t n) {
    return static_cast<T *>(std::malloc((n ? n : 1) * sizeof(T)));
}

// Checks rows [v0, v1) of an oriented CSR with row offsets of type OffT
// (offsets in range and monotone, ids strictly ascending, above their row and
// below n).  Single-threaded, so it can run inside par_chunks.  Returns
// WTC_OK or WTC_ERR_CSR.
template <class OffT>
int validate_oriented_rows_t(uint64_t n, uint64_t m, const OffT *rowPtr, const uint32_t *colInd, uint64_t v0,
                             uint64_t v1) {
    if (rowPtr[0] != 0 || (uint64_t)rowPtr[n] != m) return WTC_ERR_CSR;
    for (uint64_t v = v0; v < v1; v++) {
        const uint64_t s = rowPtr[v], t = rowPtr[v + 1];
        if (t < s || t > m) return WTC_ERR_CSR;
        uint64_t prev = v;  // every id must exceed its row and the previous id
        for (uint64_t j = s; j < t; j++) {
            const uint64_t x = colInd[j];
            if (x <= prev || x >= n) return WTC_ERR_CSR;
            prev = x;
        }
    }
    return WTC_OK;
}
int validate_oriented_rows(uint64_t n, uint64_t m, const uint64_t *rowPtr, const uint32_t *colInd, uint64_t v0,
                           uint64_t v1);

// The same check over all rows, on the host threads.
template <class OffT>
int validate_oriented_par(uint64_t n, uint64_t m, const OffT *rowPtr, const uint32_t *colInd) {
    if (!rowPtr || (m && !colInd)) return WTC_ERR_ARG;
    if (rowPtr[0] != 0 || (uint64_t)rowPtr[n] != m) return WTC_ERR_CSR;
    std::atomic<int> bad{0};
    par_chunks(n, 1 << 14, [&](uint64_t b, uint64_t e, int) {
        if (!bad.load(std::memory_order_relaxed) && validate_oriented_rows_t(n, m, rowPtr, colInd, b, e) != WTC_OK)
            bad = 1;
    });
    return bad ? WTC_ERR_CSR : WTC_OK;
}

// W = sum over rows of C(d,2) (tc.cu:348-361), over host threads.
template <class OffT>
uint64_t wedge_total_par(uint64_t n, const OffT *rowPtr) {
    std::atomic<uint64_t> W{0};
    par_chunks(n, 1 << 16, [&](uint64_t b, uint64_t e, int) {
        uint64_t w = 0;
        for (uint64_t v = b; v < e; v++) {
            const uint64_t d = (uint64_t)(rowPtr[v + 1] - rowPtr[v]);
            w += d * (d - (d > 0)) / 2;
        }
        W += w;
    });
    return W.load();
}

}  // namespace wtc

namespace wtc {
// Symmetric, deduplicated CSR of the np pairs (tok[2i] - shift,
// tok[2i+1] - shift), self-loops skipped; ids must be < n_cap.
int build_symmetric_tokens(uint64_t n_cap, const uint32_t *tok, uint64_t np, uint32_t shift,
                           wtc_big_graph **out);
}  // namespace wtc

// GPU builders of the -p 1/2 styles (wtc_prep.hip), on the process-wide
// context: symmetric CSR of loop-free pairs with ids < n, and the degree
// orientation of a symmetric CSR.  Same results as the CPU builders.
int wtc_gpu_symmetric_default(const uint32_t *pairs, uint64_t np, uint32_t n, wtc_big_graph **out);
int wtc_gpu_orient_default(const wtc_big_graph *g, wtc_big_graph **out);

namespace wtc {

static inline uint64_t binom2(uint64_t x) { return (x & 1u) ? x * ((x - 1u) >> 1) : (x >> 1) * (x - 1u); }

}  // namespace wtc
