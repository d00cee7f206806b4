// Shared helpers of the HIP kernels behind include/pyjuice_amd.h (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

// The public declarations: every PJAMD_EXPORT definition must match its declaration there.
#include "pyjuice_amd.h"

#define PJAMD_EXPORT extern "C" __attribute__((visibility("default")))

// Status codes returned by the C ABI: 0 = success, > 0 = hipError_t, < 0 = argument errors.
enum {
    PJAMD_OK = 0,
    PJAMD_EINVAL = -1,      // invalid size / unsupported combination of arguments
    PJAMD_EUNSUPPORTED = -2, // feature not implemented by this library
    PJAMD_ETOOLARGE = -3     // request exceeds the fixed device workspace of an allocation-free entry point
};

#define PJAMD_LAUNCH_CHECK()                      \
    do {                                          \
        hipError_t e_ = hipGetLastError();        \
        if (e_ != hipSuccess) return (int)e_;     \
    } while (0)

#define PJAMD_REQUIRE(cond)                       \
    do {                                          \
        if (!(cond)) return PJAMD_EINVAL;         \
    } while (0)

namespace pjamd {

constexpr float kNegInf = -INFINITY;

__device__ __forceinline__ float neg_inf() { return -INFINITY; }

// How a kernel deposits a parameter flow into its slot (float or double):
//   kPfAdd       plain read-modify-write (the slot is owned by exactly one thread of the launch)
//   kPfAtomic    atomic add (slots shared within the launch)
//   kPfOverwrite plain store: the slot is owned by one thread and the accumulator is known to be
//                zero, so it need not be read (the first backward after the flows were cleared)
enum { kPfAdd = 0, kPfAtomic = 1, kPfOverwrite = 2 };

template <typename T>
__device__ __forceinline__ void pf_add(T* p, T v, int mode) {
    if (mode == kPfAtomic) {
        atomicAdd(p, v);
    } else if (mode == kPfOverwrite) {
        *p = v;
    } else {
        *p += v;
    }
}

__device__ __forceinline__ float warp_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

template <typename T>
__device__ __forceinline__ T warp_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// Counter-based uniform in [0, 1) with 24 random bits: word 0 of Philox4x32-10 (Salmon, Moraes,
// Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC 2011) with key (k0, k1) and
// counter (c0, c1, 0, 0), mapped as (x >> 8) * 2^-24.  The sampling kernels draw the uniform of
// (node, sample) with c0 = sample, c1 = node.
__device__ __forceinline__ float philox_uniform(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1) {
    uint32_t x0 = c0, x1 = c1, x2 = 0u, x3 = 0u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, x0), lo0 = 0xD2511F53u * x0;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, x2), lo1 = 0xCD9E8D57u * x2;
        const uint32_t y0 = hi1 ^ x1 ^ k0, y2 = hi0 ^ x3 ^ k1;
        x0 = y0;
        x1 = lo1;
        x2 = y2;
        x3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return (float)(x0 >> 8) * 5.9604644775390625e-08f;
}

}  // namespace pjamd
