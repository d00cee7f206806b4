// CDNA4 32x32 matrix-core operand traits shared by the sum-layer kernels.  BF16 / FP16 operands
// use v_mfma_f32_32x32x16_{bf16,f16} (8 reduction values per lane, instruction K = 16), FP32
// operands v_mfma_f32_32x32x2_f32 (one value per lane, instruction K = 2); accumulation is FP32.
// Lane l supplies A[l & 31][KPL * (l >> 5) + q] and B[KPL * (l >> 5) + q][l & 31], q < KPL; the
// accumulator register r of lane l holds D[acc_row(r, l)][l & 31].
#pragma once

#include "pjamd_common.h"

namespace pjamd {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

template <typename T>
struct MF;

template <>
struct MF<__bf16> {
    static constexpr int KPL = 8;  // reduction values per lane per instruction (instruction K = 16)
    static constexpr int PAD = 8;  // LDS row stride of K-contiguous tiles: 16-byte aligned, 4-bank shift per row
    static __device__ __forceinline__ floatx16 mma(const __bf16* a, const __bf16* b, floatx16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(*reinterpret_cast<const bf16x8*>(a),
                                                       *reinterpret_cast<const bf16x8*>(b), c, 0, 0, 0);
    }
    static __device__ __forceinline__ __bf16 cvt(float x) { return (__bf16)x; }
};

template <>
struct MF<_Float16> {
    static constexpr int KPL = 8;
    static constexpr int PAD = 8;
    static __device__ __forceinline__ floatx16 mma(const _Float16* a, const _Float16* b, floatx16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(*reinterpret_cast<const halfx8*>(a),
                                                      *reinterpret_cast<const halfx8*>(b), c, 0, 0, 0);
    }
    static __device__ __forceinline__ _Float16 cvt(float x) { return (_Float16)x; }
};

template <>
struct MF<float> {
    static constexpr int KPL = 1;  // instruction K = 2
    static constexpr int PAD = 1;  // LDS row stride of 129 words: lanes of a column read distinct banks
    static __device__ __forceinline__ floatx16 mma(const float* a, const float* b, floatx16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(*a, *b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ float cvt(float x) { return x; }
};

// 16 x 16 x 32 matrix-core instructions (BF16 / FP16 operands, FP32 accumulation): lane l
// supplies A[l & 15][8 (l >> 4) + q] and B[8 (l >> 4) + q][l & 15], q < 8; accumulator register
// i of lane l holds D[4 (l >> 4) + i][l & 15].
typedef float floatx4 __attribute__((ext_vector_type(4)));

template <typename T>
struct MF16;

template <>
struct MF16<__bf16> {
    typedef bf16x8 vec;
    static __device__ __forceinline__ floatx4 mma(bf16x8 a, bf16x8 b, floatx4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

template <>
struct MF16<_Float16> {
    typedef halfx8 vec;
    static __device__ __forceinline__ floatx4 mma(halfx8 a, halfx8 b, floatx4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

// Matrix-core operands exp(x - s) are formed as exp2(fma(x, log2 e, -s log2 e)) with the bare
// v_exp_f32 (two VALU instructions instead of expf's denormal-range rescaling, which costs five
// more).  Results below 2^-126 flush to zero; an operand tile is scaled so that its largest entry
// is at most about 1, and BF16 / FP16 rounding of the tile's sum is far coarser than 2^-126.
// A shift s is floored at kShiftFloor, so that a column whose entries are all -inf gives
// exp2(-inf) = 0 instead of exp(-inf - (-inf)) = NaN, with no compare per element.
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kShiftFloor = -1.0e30f;

__device__ __forceinline__ float op_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

// -max(s, kShiftFloor) log2 e: the addend of exp2(fma(x, log2 e, addend)) = exp(x - s).
__device__ __forceinline__ float neg_shift2(float s) { return -fmaxf(s, kShiftFloor) * kLog2e; }

// Row of accumulator register r for this lane inside a 32 x 32 tile (the column is lane & 31).
__device__ __forceinline__ int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// log(flow / mar) of a node row entry, from flows already stored in that form (modified) or not.
__device__ __forceinline__ float lfdm_at(const float* nflows, const float* nmars, int64_t o, bool modified) {
    if (modified) return nflows[o];
    const float m = nmars[o];
    return (m == kNegInf) ? kNegInf : __logf(nflows[o]) - m;
}

}  // namespace pjamd
