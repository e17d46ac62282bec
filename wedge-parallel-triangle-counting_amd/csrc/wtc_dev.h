// wtc_dev.h -- device helpers shared by the HIP translation units: binomial
// counts, wave-uniform broadcasts and wave64/block scans.
#pragma once
#include "wtc_kernels.h"

// Launches KERNEL<uint64_t> or KERNEL<uint32_t> by the set's offset width and
// returns the launch status.
#define WTC_LAUNCH_OFF(ob, KERNEL, GRID, BLOCK, SHMEM, STREAM, ...)               \
    [&]() -> hipError_t {                                                         \
        if ((ob) == 8)                                                            \
            KERNEL<uint64_t><<<GRID, BLOCK, SHMEM, STREAM>>>(__VA_ARGS__);        \
        else                                                                      \
            KERNEL<uint32_t><<<GRID, BLOCK, SHMEM, STREAM>>>(__VA_ARGS__);        \
        return hipGetLastError();                                                 \
    }()

namespace wtc {

__device__ __forceinline__ uint64_t binom2_d(uint64_t x) {
    return (x & 1u) ? x * ((x - 1u) >> 1) : (x >> 1) * (x - 1u);
}

// C(x,2) of a 32-bit x: x(x-1) < 2^64, so one 32x32->64 multiply and a shift.
__device__ __forceinline__ uint64_t binom2_u32(uint32_t x) { return ((uint64_t)x * (uint64_t)(x - 1u)) >> 1; }

// The top-a bitmap (tc.cu:262-272, 329-346) as a square a x a bit matrix over
// the ranks [thr, thr + a), row-major in 32-bit words (a is a multiple of 32,
// so every row starts on a word boundary): the bit of (w, u) is bit
// (u - thr) % 32 of word (w - thr) * a/32 + (u - thr) / 32.  Only the strict
// upper triangle u > w is ever set or read.  a32 = a/32.
__device__ __forceinline__ uint64_t bm_word(uint32_t w, uint32_t u, uint32_t thr, uint32_t a32) {
    return (uint64_t)(w - thr) * a32 + ((u - thr) >> 5);
}
__device__ __forceinline__ uint32_t bm_test(const uint32_t *__restrict__ bitmap, uint32_t w, uint32_t u,
                                            uint32_t thr, uint32_t a32) {
    return (bitmap[bm_word(w, u, thr, a32)] >> ((u - thr) & 31u)) & 1u;
}

__device__ __forceinline__ uint32_t uni32(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// The lane-read builtins return int: each 32-bit half goes back through
// uint32_t before it is widened, so the low half is never sign-extended.
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t uni(uint32_t x) { return uni32(x); }
__device__ __forceinline__ uint64_t uni(uint64_t x) { return uni64(x); }

// Value of x in lane `src`, which must be the same in every lane.
__device__ __forceinline__ uint64_t readlane_u64(uint64_t x, uint32_t src) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, (int)src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), (int)src);
    return ((uint64_t)hi << 32) | lo;
}

// Value of x in lane src (ds_bpermute), for 64- and 32-bit values; every
// lane of the wave must execute it.
__device__ __forceinline__ uint64_t shfl_u64(uint64_t x, uint32_t src) {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)x, (int)src, kWave);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(x >> 32), (int)src, kWave);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t shfl_off(uint32_t x, uint32_t src) {
    return (uint32_t)__shfl((int)x, (int)src, kWave);
}
__device__ __forceinline__ uint64_t shfl_off(uint64_t x, uint32_t src) { return shfl_u64(x, src); }
// Value of x in lane (lane - d), or the lane's own x for lanes below d.
__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t x, int d) {
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)x, d, kWave);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(x >> 32), d, kWave);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t x) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        uint64_t y = __shfl_up(x, (unsigned)o, kWave);
        if (lane >= o) x += y;
    }
    return x;
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t x) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
    return x;
}

// Exclusive scan over the 256 lanes of a block; s_tot holds 4 wave totals.
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t x, uint64_t *s_tot, uint64_t &total) {
    const uint64_t inc = wave_incl_scan(x);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & (kWave - 1)) == kWave - 1) s_tot[w] = inc;
    __syncthreads();
    uint64_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; i++) {
        const uint64_t t = s_tot[i];
        off += (i < w) ? t : 0;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return off + inc - x;
}

// Global-memory views of the CSR, prefix and bitmap arrays.  Their pointers
// are read from descriptors in memory, so without the cast the compiler must
// treat them as generic addresses: flat loads, which also count against the
// LDS counter and make every wait cover the LDS traffic too.
template <class T>
using gptr = const __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gptr<T> to_global(const T *p) {
    return (gptr<T>)p;
}
// Constant-address-space views of the same read-only arrays, for reads whose
// address is wave-uniform: they become scalar loads into SGPRs, so the
// arithmetic that follows stays on the scalar unit.
template <class T>
using cptr = const __attribute__((address_space(4))) T *;
template <class T>
__device__ __forceinline__ cptr<T> to_const(const T *p) {
    return (cptr<T>)p;
}

// Rightmost v in [lo, hi] with ws[v] <= t, given ws[lo] <= t.  Called with
// wave-uniform arguments (scalar loads) and per lane (vector loads).
template <class P>
__device__ __forceinline__ uint32_t rightmost_le(P ws, uint32_t lo, uint32_t hi, uint64_t t) {
    uint32_t s = lo + 1, e = hi + 1;
    while (s < e) {
        const uint32_t mid = s + ((e - s) >> 1);
        if (ws[mid] <= t) s = mid + 1;
        else e = mid;
    }
    return s - 1;
}

}  // namespace wtc
