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

}  // namespace pjamd
