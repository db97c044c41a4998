// Device-side helpers shared by the katric_amd HIP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define KT_WAVE 64
#define KT_EMPTY 0xffffffffu

namespace kt {

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// Make this wave's LDS writes visible to its other lanes before they read.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    unsigned l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Inclusive prefix sum across the wave (Hillis-Steele over 6 steps).
template <typename T>
__device__ __forceinline__ T wave_incl_sum(T x) {
    unsigned l = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T y = __shfl_up(x, d, 64);
        if (l >= (unsigned)d) x += y;
    }
    return x;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// Four-bit degree class, monotone in the degree: floor(log2 d) capped at 15
// (0 for d < 2).  Two classes per byte, so the class table of n vertices is
// n / 2 bytes and stays in the L2 beside the streamed heads; equal classes
// are decided by the exact degrees.
__device__ __forceinline__ uint32_t deg_class4(uint32_t d) {
    return d < 2 ? 0u : min(15u, 31u - (uint32_t)__clz(d));
}

__device__ __forceinline__ uint32_t class4_of(const uint8_t* __restrict__ ncls, uint32_t u) {
    return (ncls[u >> 1] >> ((u & 1) * 4)) & 15u;
}

// 64-bit finaliser of MurmurHash3 (fmix64): a bijective mix in which every
// input bit affects every output bit.
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

// Signed contribution of the slot v -> u to a keyed checksum over unordered
// pairs: +h({v,u}) when v < u, -h({v,u}) when v > u, so the two slots of an
// undirected edge cancel.  h is two rounds of 32-bit multiply-xorshift
// (murmur-style), the second round seeded by the first, so both halves of the
// 64-bit value depend on both ids and on the key; two quarter-rate multiplies
// per slot (the orientation pass that computes it is VALU-bound).
__device__ __forceinline__ uint64_t pair_term(uint32_t v, uint32_t u, uint64_t key) {
    const uint32_t a = min(v, u), b = max(v, u);
    uint32_t x = (a ^ (uint32_t)key) * 0x9E3779B1u;
    x ^= x >> 15;
    uint32_t y = ((b ^ (uint32_t)(key >> 32)) + x) * 0x85EBCA77u;
    y ^= y >> 13;
    const uint64_t h = ((uint64_t)(x ^ __builtin_amdgcn_alignbit(y, y, 11)) << 32) | y;
    return v < u ? h : (0ull - h);
}

// Graph g in [glo, ghi] with vbase[g] <= v < vbase[g + 1] (the last g with
// vbase[g] <= v, so that empty graphs sharing v's base are skipped).
__device__ __forceinline__ uint32_t graph_of(const uint64_t* __restrict__ vbase, uint32_t glo, uint32_t ghi,
                                             uint64_t v) {
    while (glo < ghi) {
        const uint32_t mid = glo + (ghi - glo + 1) / 2;
        if (vbase[mid] <= v)
            glo = mid;
        else
            ghi = mid - 1;
    }
    return glo;
}

// ORs flag bits into a word many threads report to, skipping the atomic when
// the bits are set already (atomics on one word serialise in L2).
__device__ __forceinline__ void or_flags(uint32_t* __restrict__ err, uint32_t flags) {
    if (flags && (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & flags) != flags)
        atomicOr(err, flags);
}

__device__ __forceinline__ unsigned long long atomic_add_u64(uint64_t* p, uint64_t v) {
    return atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v));
}

// Largest v in [lo, hi] with off[v] <= s (the CSR row holding slot s);
// requires off[lo] <= s.
__device__ __forceinline__ uint32_t row_of_slot(const uint64_t* __restrict__ off, uint32_t lo, uint32_t hi,
                                                uint64_t s) {
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo + 1) / 2;
        if (off[mid] <= s)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint32_t wave_incl_max_u32(uint32_t x) {
    unsigned l = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (l >= (unsigned)d) x = max(x, y);
    }
    return x;
}

// Resolves the CSR row of every slot of the tile [tile0, tend) of a CSR with
// n rows: afterwards marks[i] = row(tile0 + i) - lo for i < tend - tile0.
// Rows starting strictly inside the tile mark their first slot (the largest
// row wins when empty rows share a start), then an inclusive max-scan over
// the THREADS * ITEMS marks fills the gaps.  The rows of the tile's first and
// last slot come from `given` (two entries per tile, from k_tile_bounds) or,
// when it is null, from two binary searches over the offsets.  Block-wide:
// every thread calls it.
template <int THREADS, int ITEMS>
__device__ __forceinline__ void tile_rows(const uint64_t* __restrict__ off, uint32_t n, uint64_t tile0, uint64_t tend,
                                          uint32_t* marks, uint32_t* bnd, uint32_t* wmax, uint32_t& lo,
                                          uint32_t& hi, const uint32_t* __restrict__ given = nullptr) {
    constexpr int TILE = THREADS * ITEMS;
    const unsigned tid = threadIdx.x;
    if (tid == 0) {
        if (given) {  // clamped: offsets that fail validation must not send the scan out of range
            const uint64_t t = tile0 / TILE;
            const uint32_t a = min(given[2 * t], n - 1), b = min(given[2 * t + 1], n - 1);
            bnd[0] = a;
            bnd[1] = max(a, b);
        } else {
            bnd[0] = row_of_slot(off, 0, n - 1, tile0);
            bnd[1] = row_of_slot(off, 0, n - 1, tend - 1);
        }
    }
    for (int i = tid; i < TILE; i += THREADS) marks[i] = 0;
    __syncthreads();
    lo = bnd[0];
    hi = bnd[1];
    for (uint64_t r = (uint64_t)lo + 1 + tid; r <= hi; r += THREADS) {
        uint64_t pos = off[r] - tile0;
        if (pos < (uint64_t)TILE) atomicMax(&marks[pos], (uint32_t)(r - lo));
    }
    __syncthreads();
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) m = max(m, marks[tid * ITEMS + i]);
    uint32_t incl = wave_incl_max_u32(m);
    if (lane_id() == KT_WAVE - 1) wmax[tid / KT_WAVE] = incl;
    __syncthreads();
    uint32_t carry = __shfl_up(incl, 1, 64);
    if (lane_id() == 0) carry = 0;
    for (unsigned w = 0; w < tid / KT_WAVE; ++w) carry = max(carry, wmax[w]);
    uint32_t run = carry;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        run = max(run, marks[tid * ITEMS + i]);
        marks[tid * ITEMS + i] = run;
    }
    __syncthreads();
}

// Sum over the wave with DPP row shifts and row broadcasts (no LDS traffic):
// an inclusive scan within each 16-lane row by shifts of 1, 2, 4, 8 (lanes
// shifted in from outside the row read 0), then lane 15 of rows 0 and 2 into
// rows 1 and 3, then lane 31 into rows 2 and 3; lane 63 holds the total.
__device__ __forceinline__ uint32_t wave_total_dpp(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false); // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false); // row_bcast:31
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

}  // namespace kt
