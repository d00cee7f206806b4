// Sum layers on MFMA (CDNA4 32x32 matrix cores): forward log-sum-exp, element flows and
// parameter flows for block sizes >= 32.  Operands are BF16, FP16 or FP32 (v_mfma_f32_32x32x16_bf16,
// v_mfma_f32_32x32x16_f16, v_mfma_f32_32x32x2_f32); accumulation is always FP32.
//
// Every kernel is one workgroup of 4 wavefronts computing a TM x TN tile (each wavefront one
// 32 x 32 MFMA accumulator).  Operand tiles are staged through LDS with K (the reduction index)
// contiguous, TK = 32 reduction steps per stage.  Range is controlled by max subtraction:
//   forward   : exp(e - colmax_over_children)          (values in [0, 1])
//   ele-flows : exp(log(f / m) - colmax_over_parents)  (values in [0, 1])
//   par-flows : exp(log(f / m) - s_b) * exp(e + s_b),  s_b = max over the tile's parent rows
#include "pjamd_common.h"

using namespace pjamd;

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

constexpr int TK = 32;

template <typename T>
struct MF;

template <>
struct MF<__bf16> {
    static constexpr int KPL = 8;  // reduction values per lane per instruction (instruction K = 16)
    static __device__ __forceinline__ floatx16 mma(const __bf16* a, const __bf16* b, floatx16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(*reinterpret_cast<const bf16x8*>(a),
                                                       *reinterpret_cast<const bf16x8*>(b), c, 0, 0, 0);
    }
    static __device__ __forceinline__ __bf16 cvt(float x) { return (__bf16)x; }
};

template <>
struct MF<_Float16> {
    static constexpr int KPL = 8;
    static __device__ __forceinline__ floatx16 mma(const _Float16* a, const _Float16* b, floatx16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(*reinterpret_cast<const halfx8*>(a),
                                                      *reinterpret_cast<const halfx8*>(b), c, 0, 0, 0);
    }
    static __device__ __forceinline__ _Float16 cvt(float x) { return (_Float16)x; }
};

template <>
struct MF<float> {
    static constexpr int KPL = 1;  // instruction K = 2
    static __device__ __forceinline__ floatx16 mma(const float* a, const float* b, floatx16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(*a, *b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ float cvt(float x) { return x; }
};

// Row of accumulator register r for this lane inside a 32 x 32 tile (the column is lane & 31).
__device__ __forceinline__ int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

template <typename T, int WM, int WN, int LDA, int LDB>
__device__ __forceinline__ floatx16 stage_mma(const T (*As)[LDA], const T (*Bs)[LDB], int wm, int wn, int lane,
                                              floatx16 acc) {
#pragma unroll
    for (int kk = 0; kk < TK; kk += 2 * MF<T>::KPL) {
        acc = MF<T>::mma(&As[wm * 32 + (lane & 31)][kk + MF<T>::KPL * (lane >> 5)],
                         &Bs[wn * 32 + (lane & 31)][kk + MF<T>::KPL * (lane >> 5)], acc);
    }
    return acc;
}

__device__ __forceinline__ float lfdm_at(const float* nflows, const float* nmars, int64_t o, bool modified) {
    if (modified) return nflows[o];
    const float m = nmars[o];
    return (m == kNegInf) ? kNegInf : __logf(nflows[o]) - m;
}

// Column max of TN batch columns over `rows` rows given by a row-id functor.
template <int TN, typename RowFn>
__device__ __forceinline__ void column_max(float* cmax, float (*red)[TN], int rows, int64_t b0, int64_t B, RowFn rowval) {
    constexpr int G = 256 / TN;
    const int tid = threadIdx.x;
    const int col = tid % TN, g = tid / TN;
    const int64_t b = b0 + col;
    float m = kNegInf;
    if (b < B)
        for (int k = g; k < rows; k += G) m = fmaxf(m, rowval(k, b));
    red[g][col] = m;
    __syncthreads();
    if (tid < TN) {
        float t = red[0][tid];
#pragma unroll
        for (int q = 1; q < G; ++q) t = fmaxf(t, red[q][tid]);
        cmax[tid] = t;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------------------------
template <typename T, int TM, int TN>
__global__ __launch_bounds__(256) void sum_fwd_mfma(float* __restrict__ node_mars, const float* __restrict__ emars,
                                                    const float* __restrict__ params, const int* __restrict__ nids,
                                                    const int* __restrict__ cids, const int* __restrict__ pids,
                                                    const int* __restrict__ nch, int K, int bs, int64_t B) {
    constexpr int WM = TM / 32, WN = TN / 32;
    static_assert(WM * WN == 4, "four wavefronts per workgroup");
    constexpr int PAD = 16 / sizeof(T);
    constexpr int LD = TK + PAD;
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    __shared__ float cmax[TN];
    __shared__ float red[256 / TN][TN];

    const int mtiles = bs / TM;
    const int j = blockIdx.x / mtiles, m0 = (blockIdx.x % mtiles) * TM;
    const int64_t b0 = (int64_t)blockIdx.y * TN;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w % WM, wn = w / WM;
    const int n = nch[j];
    const int* crow = cids + (int64_t)j * K;
    const int* prow = pids + (int64_t)j * K;

    column_max<TN>(cmax, red, n, b0, B, [&](int k, int64_t b) { return emars[(int64_t)crow[k] * B + b]; });

    constexpr int PB = TK * TN / 256;
    const int bcol = tid % TN, bk0 = (tid / TN) * PB;
    const float bm = cmax[bcol];
    const int64_t bb = b0 + bcol;
    constexpr int PA = TK * TM / 256;
    const int arow = tid % TM, ak0 = (tid / TM) * PA;

    floatx16 acc = {0};
    for (int k0 = 0; k0 < n; k0 += TK) {
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int k = k0 + bk0 + q;
            float x = 0.0f;
            if (k < n && bb < B && bm != kNegInf) x = __expf(emars[(int64_t)crow[k] * B + bb] - bm);
            Bs[bcol][bk0 + q] = MF<T>::cvt(x);
        }
#pragma unroll
        for (int q = 0; q < PA; ++q) {
            const int k = k0 + ak0 + q;
            As[arow][ak0 + q] = MF<T>::cvt(k < n ? params[(int64_t)prow[k] + m0 + arow] : 0.0f);
        }
        __syncthreads();
        acc = stage_mma<T, WM, WN, LD, LD>(As, Bs, wm, wn, lane, acc);
        __syncthreads();
    }
    const int col = wn * 32 + (lane & 31);
    const int64_t b = b0 + col;
    if (b < B) {
        const float m = cmax[col];
        const int64_t base = (int64_t)nids[j] + m0 + wm * 32;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float v = acc[r];
            node_mars[(base + acc_row(r, lane)) * B + b] = (v > 0.0f && m != kNegInf) ? __logf(v) + m : kNegInf;
        }
    }
}

// ---------------------------------------------------------------------------------------------
template <typename T, int TM, int TN>
__global__ __launch_bounds__(256) void sum_ele_mfma(float* __restrict__ eflows, const float* __restrict__ nflows,
                                                    const float* __restrict__ nmars, const float* __restrict__ emars,
                                                    const float* __restrict__ params, const int* __restrict__ chids,
                                                    const int* __restrict__ parids, const int* __restrict__ parpids,
                                                    const int* __restrict__ npar, int KP, int cbs, int bs, int64_t B,
                                                    bool modified, bool accumulate) {
    constexpr int WM = TM / 32, WN = TN / 32;
    static_assert(WM * WN == 4, "four wavefronts per workgroup");
    constexpr int PAD = 16 / sizeof(T);
    constexpr int LD = TK + PAD;
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    __shared__ float cmax[TN];
    __shared__ float red[256 / TN][TN];

    const int mtiles = cbs / TM;
    const int j = blockIdx.x / mtiles, m0 = (blockIdx.x % mtiles) * TM;
    const int64_t b0 = (int64_t)blockIdx.y * TN;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w % WM, wn = w / WM;
    const int kflat = npar[j] * bs;
    const int* parrow = parids + (int64_t)j * KP;
    const int* pprow = parpids + (int64_t)j * KP;

    column_max<TN>(cmax, red, kflat, b0, B, [&](int kk, int64_t b) {
        const int kb = kk / bs, pi = kk - kb * bs;
        return lfdm_at(nflows, nmars, ((int64_t)parrow[kb] + pi) * B + b, modified);
    });

    constexpr int PB = TK * TN / 256;
    const int bcol = tid % TN, bk0 = (tid / TN) * PB;
    const float bm = cmax[bcol];
    const int64_t bb = b0 + bcol;
    constexpr int PA = TK * TM / 256;
    const int arow = tid % TM, ak0 = (tid / TM) * PA;

    floatx16 acc = {0};
    for (int k0 = 0; k0 < kflat; k0 += TK) {
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int kk = k0 + bk0 + q;
            float x = 0.0f;
            if (kk < kflat && bb < B && bm != kNegInf) {
                const int kb = kk / bs, pi = kk - kb * bs;
                const float l = lfdm_at(nflows, nmars, ((int64_t)parrow[kb] + pi) * B + bb, modified);
                x = __expf(l - bm);
            }
            Bs[bcol][bk0 + q] = MF<T>::cvt(x);
        }
#pragma unroll
        for (int q = 0; q < PA; ++q) {
            const int kk = k0 + ak0 + q;
            float a = 0.0f;
            if (kk < kflat) {
                const int kb = kk / bs, pi = kk - kb * bs;
                a = params[(int64_t)pprow[kb] + (int64_t)(m0 + arow) * bs + pi];
            }
            As[arow][ak0 + q] = MF<T>::cvt(a);
        }
        __syncthreads();
        acc = stage_mma<T, WM, WN, LD, LD>(As, Bs, wm, wn, lane, acc);
        __syncthreads();
    }
    const int col = wn * 32 + (lane & 31);
    const int64_t b = b0 + col;
    if (b < B) {
        const float m = cmax[col];
        const int64_t base = (int64_t)chids[j] + m0 + wm * 32;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t o = (base + acc_row(r, lane)) * B + b;
            const float e = emars[o];
            float v = (m == kNegInf || e == kNegInf) ? 0.0f : acc[r] * __expf(e + m);
            if (accumulate) v += eflows[o];
            eflows[o] = v;
        }
    }
}

// ---------------------------------------------------------------------------------------------
template <typename T, int TM, int TN, typename PF>
__global__ __launch_bounds__(256) void sum_par_mfma(PF* __restrict__ pflows, const float* __restrict__ nflows,
                                                    const float* __restrict__ nmars, const float* __restrict__ emars,
                                                    const float* __restrict__ params, const int* __restrict__ nids,
                                                    const int* __restrict__ cids, const int* __restrict__ pids,
                                                    const int* __restrict__ pfids, const int* __restrict__ nch, int K,
                                                    int bs, int64_t B, int64_t bchunk, bool modified, bool atomic_out,
                                                    float sign) {
    constexpr int WM = TM / 32, WN = TN / 32;
    static_assert(WM * WN == 4, "four wavefronts per workgroup");
    constexpr int PAD = 16 / sizeof(T);
    constexpr int LD = TK + PAD;
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    __shared__ float Xr[TM][TK + 1];
    __shared__ float sh[TK];

    const int mtiles = bs / TM;
    const int j = blockIdx.x / mtiles, m0 = (blockIdx.x % mtiles) * TM;
    const int kt0 = blockIdx.y * TN;
    const int n = nch[j];
    if (kt0 >= n) return;
    const int64_t bstart = (int64_t)blockIdx.z * bchunk;
    const int64_t bend = (bstart + bchunk < B) ? bstart + bchunk : B;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w % WM, wn = w / WM;
    const int64_t nrow0 = (int64_t)nids[j] + m0;
    const int* crow = cids + (int64_t)j * K;

    constexpr int PX = TK * TM / 256;   // consecutive batch columns per thread, X tile
    constexpr int XT = TK / PX;         // threads per row
    const int xrow = tid / XT, xb0 = (tid % XT) * PX;
    constexpr int PY = TK * TN / 256;
    constexpr int YT = TK / PY;
    const int yrow = tid / YT, yb0 = (tid % YT) * PY;
    const int yk = kt0 + yrow;
    const int64_t ybase = (yk < n) ? (int64_t)crow[yk] * B : -1;

    floatx16 acc = {0};
    for (int64_t b0 = bstart; b0 < bend; b0 += TK) {
        float xv[PX];
#pragma unroll
        for (int q = 0; q < PX; ++q) {
            const int64_t b = b0 + xb0 + q;
            xv[q] = (b < bend) ? lfdm_at(nflows, nmars, (nrow0 + xrow) * B + b, modified) : kNegInf;
            Xr[xrow][xb0 + q] = xv[q];
        }
        __syncthreads();
        if (tid < TK) {
            float t = kNegInf;
            for (int r = 0; r < TM; ++r) t = fmaxf(t, Xr[r][tid]);
            sh[tid] = t;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PX; ++q) {
            const float s = sh[xb0 + q];
            As[xrow][xb0 + q] = MF<T>::cvt((s == kNegInf || xv[q] == kNegInf) ? 0.0f : __expf(xv[q] - s));
        }
#pragma unroll
        for (int q = 0; q < PY; ++q) {
            const int64_t b = b0 + yb0 + q;
            const float s = sh[yb0 + q];
            float y = 0.0f;
            if (ybase >= 0 && b < bend && s != kNegInf) {
                const float e = emars[ybase + b];
                y = (e == kNegInf) ? 0.0f : __expf(e + s);
            }
            Bs[yrow][yb0 + q] = MF<T>::cvt(y);
        }
        __syncthreads();
        acc = stage_mma<T, WM, WN, LD, LD>(As, Bs, wm, wn, lane, acc);
        __syncthreads();
    }
    const int k = kt0 + wn * 32 + (lane & 31);
    if (k < n) {
        const int64_t p0 = (int64_t)pids[(int64_t)j * K + k] + m0 + wm * 32;
        const int64_t f0 = (int64_t)pfids[(int64_t)j * K + k] + m0 + wm * 32;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = acc_row(r, lane);
            const float v = sign * params[p0 + i] * acc[r];
            pf_add<PF>(&pflows[f0 + i], (PF)v, atomic_out);
        }
    }
}

// One 32 x 32 x KI product with the operand / accumulator lane layouts the kernels above assume:
// A is [32][KI] row-major, B is [KI][32] row-major, C is [32][32] row-major.
template <typename T>
__global__ void mfma_layout_probe(const float* A, const float* B, float* C) {
    constexpr int KI = 2 * MF<T>::KPL;
    __shared__ __attribute__((aligned(16))) T As[32][KI];
    __shared__ __attribute__((aligned(16))) T Bs[32][KI];
    const int lane = threadIdx.x;
    for (int t = lane; t < 32 * KI; t += 64) {
        const int r = t / KI, k = t % KI;
        As[r][k] = MF<T>::cvt(A[r * KI + k]);
        Bs[r][k] = MF<T>::cvt(B[k * 32 + r]);
    }
    __syncthreads();
    floatx16 acc = {0};
    acc = MF<T>::mma(&As[lane & 31][MF<T>::KPL * (lane >> 5)], &Bs[lane & 31][MF<T>::KPL * (lane >> 5)], acc);
    for (int r = 0; r < 16; ++r) C[acc_row(r, lane) * 32 + (lane & 31)] = acc[r];
}

template <typename T>
int launch_fwd(float* node_mars, const float* emars, const float* params, const int* nids, const int* cids,
               const int* pids, const int* nch, int n_blocks, int K, int bs, int64_t B, hipStream_t stream) {
    if (bs >= 64) {
        dim3 grid((unsigned)((int64_t)n_blocks * (bs / 64)), cdiv(B, 64));
        sum_fwd_mfma<T, 64, 64><<<grid, 256, 0, stream>>>(node_mars, emars, params, nids, cids, pids, nch, K, bs, B);
    } else {
        dim3 grid((unsigned)n_blocks, cdiv(B, 128));
        sum_fwd_mfma<T, 32, 128><<<grid, 256, 0, stream>>>(node_mars, emars, params, nids, cids, pids, nch, K, bs, B);
    }
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

template <typename T>
int launch_ele(float* eflows, const float* nflows, const float* nmars, const float* emars, const float* params,
               const int* chids, const int* parids, const int* parpids, const int* npar, int n, int KP, int cbs, int bs,
               int64_t B, bool mod, bool accum, hipStream_t stream) {
    if (cbs >= 64) {
        dim3 grid((unsigned)((int64_t)n * (cbs / 64)), cdiv(B, 64));
        sum_ele_mfma<T, 64, 64><<<grid, 256, 0, stream>>>(eflows, nflows, nmars, emars, params, chids, parids, parpids,
                                                         npar, KP, cbs, bs, B, mod, accum);
    } else {
        dim3 grid((unsigned)n, cdiv(B, 128));
        sum_ele_mfma<T, 32, 128><<<grid, 256, 0, stream>>>(eflows, nflows, nmars, emars, params, chids, parids, parpids,
                                                          npar, KP, cbs, bs, B, mod, accum);
    }
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

template <typename T, typename PF>
int launch_par(PF* pflows, const float* nflows, const float* nmars, const float* emars, const float* params,
               const int* nids, const int* cids, const int* pids, const int* pfids, const int* nch, int n_blocks, int K,
               int bs, int64_t B, bool mod, bool shared, float sign, hipStream_t stream) {
    const int TM = bs >= 64 ? 64 : 32;
    const int TN = bs >= 64 ? 64 : 128;
    const int64_t tiles = (int64_t)n_blocks * (bs / TM) * ((K + TN - 1) / TN);
    // Split the batch so that the launch has >= ~2048 workgroups; partial sums are added atomically.
    int64_t splits = (2048 + tiles - 1) / tiles;
    const int64_t max_splits = (B + TK - 1) / TK;
    if (splits > max_splits) splits = max_splits;
    if (splits > 16) splits = 16;
    if (splits < 1) splits = 1;
    int64_t bchunk = (B + splits - 1) / splits;
    bchunk = ((bchunk + TK - 1) / TK) * TK;
    splits = (B + bchunk - 1) / bchunk;
    const bool atomic_out = shared || splits > 1;
    dim3 grid((unsigned)((int64_t)n_blocks * (bs / TM)), (unsigned)((K + TN - 1) / TN), (unsigned)splits);
    if (bs >= 64)
        sum_par_mfma<T, 64, 64, PF><<<grid, 256, 0, stream>>>(pflows, nflows, nmars, emars, params, nids, cids, pids,
                                                             pfids, nch, K, bs, B, bchunk, mod, atomic_out, sign);
    else
        sum_par_mfma<T, 32, 128, PF><<<grid, 256, 0, stream>>>(pflows, nflows, nmars, emars, params, nids, cids, pids,
                                                              pfids, nch, K, bs, B, bchunk, mod, atomic_out, sign);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

}  // namespace

// precision: 0 = FP32 operands, 1 = BF16, 2 = FP16 (FP32 accumulation in every case).
int pjamd_sum_forward_mfma(float* node_mars, const float* element_mars, const float* params, const int* nids,
                           const int* cids, const int* pids, const int* nch, int n_blocks, int K, int bs, int64_t B,
                           int precision, hipStream_t stream) {
    if (bs < 32) return PJAMD_EUNSUPPORTED;
    if (precision == 1)
        return launch_fwd<__bf16>(node_mars, element_mars, params, nids, cids, pids, nch, n_blocks, K, bs, B, stream);
    if (precision == 2)
        return launch_fwd<_Float16>(node_mars, element_mars, params, nids, cids, pids, nch, n_blocks, K, bs, B, stream);
    return launch_fwd<float>(node_mars, element_mars, params, nids, cids, pids, nch, n_blocks, K, bs, B, stream);
}

int pjamd_sum_backward_ele_mfma(float* eflows, const float* nflows, const float* nmars, const float* emars,
                                const float* params, const int* chids, const int* parids, const int* parpids,
                                const int* npar, int n, int KP, int cbs, int bs, int64_t B, bool mod, bool accum,
                                int precision, hipStream_t stream) {
    if (cbs < 32) return PJAMD_EUNSUPPORTED;
    if (precision == 1)
        return launch_ele<__bf16>(eflows, nflows, nmars, emars, params, chids, parids, parpids, npar, n, KP, cbs, bs, B,
                                  mod, accum, stream);
    if (precision == 2)
        return launch_ele<_Float16>(eflows, nflows, nmars, emars, params, chids, parids, parpids, npar, n, KP, cbs, bs,
                                    B, mod, accum, stream);
    return launch_ele<float>(eflows, nflows, nmars, emars, params, chids, parids, parpids, npar, n, KP, cbs, bs, B,
                             mod, accum, stream);
}

// FP16 operands can overflow in the parameter-flow product (exp(e + s) is unbounded above), so
// precision 2 runs this contraction in BF16.
int pjamd_sum_backward_par_mfma(void* pflows, bool pf64, const float* nflows, const float* nmars, const float* emars,
                                const float* params, const int* nids, const int* cids, const int* pids, const int* pfids,
                                const int* nch, int n_blocks, int K, int bs, int64_t B, bool mod, bool shared, float sign,
                                int precision, hipStream_t stream) {
    if (bs < 32) return PJAMD_EUNSUPPORTED;
    if (precision == 0) {
        if (pf64)
            return launch_par<float, double>((double*)pflows, nflows, nmars, emars, params, nids, cids, pids, pfids, nch,
                                             n_blocks, K, bs, B, mod, shared, sign, stream);
        return launch_par<float, float>((float*)pflows, nflows, nmars, emars, params, nids, cids, pids, pfids, nch,
                                        n_blocks, K, bs, B, mod, shared, sign, stream);
    }
    if (pf64)
        return launch_par<__bf16, double>((double*)pflows, nflows, nmars, emars, params, nids, cids, pids, pfids, nch,
                                          n_blocks, K, bs, B, mod, shared, sign, stream);
    return launch_par<__bf16, float>((float*)pflows, nflows, nmars, emars, params, nids, cids, pids, pfids, nch,
                                     n_blocks, K, bs, B, mod, shared, sign, stream);
}

// Layout probe used by the tests: C = A (32 x KI) * B (KI x 32) with KI = 16 (bf16 / fp16) or 2 (fp32).
PJAMD_EXPORT int pjamd_debug_mfma_probe(int precision, const float* A, const float* B, float* C, hipStream_t stream) {
    if (precision == 1) mfma_layout_probe<__bf16><<<1, 64, 0, stream>>>(A, B, C);
    else if (precision == 2) mfma_layout_probe<_Float16><<<1, 64, 0, stream>>>(A, B, C);
    else if (precision == 0) mfma_layout_probe<float><<<1, 64, 0, stream>>>(A, B, C);
    else return PJAMD_EINVAL;
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}
