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

constexpr int KCAP = 1024;  // child / parent index entries staged in LDS per chunk

// ---------------------------------------------------------------------------------------------
// Forward.  Workgroup = (node block j, TM parent rows) x TN batch columns.  The child index rows
// are staged in LDS; the next stage's operands are prefetched into registers while the matrix
// cores work on the current one; the column max is a running max (accumulators are rescaled).
template <typename T, int TM, int TN>
__global__ __launch_bounds__(256) void sum_fwd_mfma(float* __restrict__ node_mars, const float* __restrict__ emars,
                                                    const float* __restrict__ params, const int* __restrict__ nids,
                                                    const int* __restrict__ cids, const int* __restrict__ pids,
                                                    const int* __restrict__ nch, int K, int bs, int64_t B) {
    constexpr int WM = TM / 32, WN = TN / 32;
    static_assert(WM * WN == 4, "four wavefronts per workgroup");
    constexpr int PAD = 16 / sizeof(T);
    constexpr int LD = TK + PAD;
    constexpr int G = 256 / TN, PB = TK / G;       // element tile: G row groups x TN columns
    constexpr int PA = TK * TM / 256;              // parameter tile: PA consecutive k per thread
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    __shared__ int sc[KCAP], sp[KCAP];
    __shared__ float red[G][TN], mrun[TN], msc[TN];

    const int mtiles = bs / TM;
    const int j = blockIdx.x / mtiles, m0 = (blockIdx.x % mtiles) * TM;
    const int64_t b0 = (int64_t)blockIdx.y * TN;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w % WM, wn = w / WM;
    const int n = nch[j];
    const int* crow = cids + (int64_t)j * K;
    const int* prow = pids + (int64_t)j * K;
    const int bcol = tid % TN, bg = tid / TN, bk0 = bg * PB;
    const int64_t bb = b0 + bcol;
    const bool bok = bb < B;
    const int arow = tid % TM, ak0 = (tid / TM) * PA;
    const int mycol = wn * 32 + (lane & 31);

    if (tid < TN) mrun[tid] = kNegInf;
    floatx16 acc = {0};
    float ev[PB], av[PA];
    for (int cb = 0; cb < n; cb += KCAP) {
        const int cn = (n - cb < KCAP) ? n - cb : KCAP;
        __syncthreads();
        for (int t = tid; t < cn; t += 256) {
            sc[t] = crow[cb + t];
            sp[t] = prow[cb + t];
        }
        __syncthreads();
        auto load = [&](int k0) {
#pragma unroll
            for (int q = 0; q < PB; ++q) {
                const int k = k0 + bk0 + q;
                ev[q] = (k < cn && bok) ? emars[(int64_t)sc[k] * B + bb] : kNegInf;
            }
#pragma unroll
            for (int q = 0; q < PA; ++q) {
                const int k = k0 + ak0 + q;
                av[q] = (k < cn) ? params[(int64_t)sp[k] + m0 + arow] : 0.0f;
            }
        };
        load(0);
        for (int k0 = 0; k0 < cn; k0 += TK) {
            float lm = kNegInf;
#pragma unroll
            for (int q = 0; q < PB; ++q) lm = fmaxf(lm, ev[q]);
            red[bg][bcol] = lm;
            __syncthreads();
            if (tid < TN) {
                float t = red[0][tid];
#pragma unroll
                for (int g = 1; g < G; ++g) t = fmaxf(t, red[g][tid]);
                const float mo = mrun[tid];
                const float mn = fmaxf(mo, t);
                msc[tid] = (mn == kNegInf) ? 1.0f : __expf(mo - mn);
                mrun[tid] = mn;
            }
            __syncthreads();
            const float mn = mrun[bcol];
#pragma unroll
            for (int q = 0; q < PB; ++q) Bs[bcol][bk0 + q] = MF<T>::cvt(mn == kNegInf ? 0.0f : __expf(ev[q] - mn));
#pragma unroll
            for (int q = 0; q < PA; ++q) As[arow][ak0 + q] = MF<T>::cvt(av[q]);
            const float s = msc[mycol];
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] *= s;
            __syncthreads();
            if (k0 + TK < cn) load(k0 + TK);
            acc = stage_mma<T, WM, WN, LD, LD>(As, Bs, wm, wn, lane, acc);
        }
    }
    const int64_t b = b0 + mycol;
    if (b < B) {
        const float m = mrun[mycol];
        const int64_t base = (int64_t)nids[j] + m0 + wm * 32;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float v = acc[r];
            node_mars[(base + acc_row(r, lane)) * B + b] = (v > 0.0f && m != kNegInf) ? __logf(v) + m : kNegInf;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Element flows.  Workgroup = (child block j, TM child rows) x TN batch columns; the reduction
// runs over the flattened (parent block, parent row) index with a running max of log(f / m).
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
    constexpr int G = 256 / TN, PB = TK / G;
    constexpr int PA = TK * TM / 256;
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    __shared__ int spar[KCAP], spp[KCAP];
    __shared__ float red[G][TN], mrun[TN], msc[TN];

    const int mtiles = cbs / TM;
    const int j = blockIdx.x / mtiles, m0 = (blockIdx.x % mtiles) * TM;
    const int64_t b0 = (int64_t)blockIdx.y * TN;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w % WM, wn = w / WM;
    const int np = npar[j];
    const int* parrow = parids + (int64_t)j * KP;
    const int* pprow = parpids + (int64_t)j * KP;
    const int bcol = tid % TN, bg = tid / TN, bk0 = bg * PB;
    const int64_t bb = b0 + bcol;
    const bool bok = bb < B;
    // parameter tile: a child row's parameters are contiguous along the parent index, so lanes
    // take PA consecutive reduction indices of one row
    constexpr int AT = TK / PA;                    // threads per row
    const int arow = tid / AT, ak0 = (tid % AT) * PA;
    const int mycol = wn * 32 + (lane & 31);
    const int64_t arow_off = (int64_t)(m0 + arow) * bs;

    if (tid < TN) mrun[tid] = kNegInf;
    floatx16 acc = {0};
    float lv[PB], av[PA];
    for (int pb = 0; pb < np; pb += KCAP) {
        const int cnp = (np - pb < KCAP) ? np - pb : KCAP;
        const int kf = cnp * bs;
        __syncthreads();
        for (int t = tid; t < cnp; t += 256) {
            spar[t] = parrow[pb + t];
            spp[t] = pprow[pb + t];
        }
        __syncthreads();
        auto load = [&](int k0) {
#pragma unroll
            for (int q = 0; q < PB; ++q) {
                const int kk = k0 + bk0 + q;
                float v = kNegInf;
                if (kk < kf && bok) {
                    const int kb = kk / bs, pi = kk - kb * bs;
                    v = lfdm_at(nflows, nmars, ((int64_t)spar[kb] + pi) * B + bb, modified);
                }
                lv[q] = v;
            }
#pragma unroll
            for (int q = 0; q < PA; ++q) {
                const int kk = k0 + ak0 + q;
                float a = 0.0f;
                if (kk < kf) {
                    const int kb = kk / bs, pi = kk - kb * bs;
                    a = params[(int64_t)spp[kb] + arow_off + pi];
                }
                av[q] = a;
            }
        };
        load(0);
        for (int k0 = 0; k0 < kf; k0 += TK) {
            float lm = kNegInf;
#pragma unroll
            for (int q = 0; q < PB; ++q) lm = fmaxf(lm, lv[q]);
            red[bg][bcol] = lm;
            __syncthreads();
            if (tid < TN) {
                float t = red[0][tid];
#pragma unroll
                for (int g = 1; g < G; ++g) t = fmaxf(t, red[g][tid]);
                const float mo = mrun[tid];
                const float mn = fmaxf(mo, t);
                msc[tid] = (mn == kNegInf) ? 1.0f : __expf(mo - mn);
                mrun[tid] = mn;
            }
            __syncthreads();
            const float mn = mrun[bcol];
#pragma unroll
            for (int q = 0; q < PB; ++q) Bs[bcol][bk0 + q] = MF<T>::cvt(mn == kNegInf ? 0.0f : __expf(lv[q] - mn));
#pragma unroll
            for (int q = 0; q < PA; ++q) As[arow][ak0 + q] = MF<T>::cvt(av[q]);
            const float s = msc[mycol];
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] *= s;
            __syncthreads();
            if (k0 + TK < kf) load(k0 + TK);
            acc = stage_mma<T, WM, WN, LD, LD>(As, Bs, wm, wn, lane, acc);
        }
    }
    const int64_t b = b0 + mycol;
    if (b < B) {
        const float m = mrun[mycol];
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
// Parameter flows.  Workgroup = (node block j, TM parent rows) x (TN children), reduction over
// the batch range [blockIdx.z * bchunk, ...).  Operands of the next stage are prefetched.
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
    __shared__ float xred[256 / TK][TK];
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

    constexpr int XG = 256 / TK;        // row groups of the X tile; lanes run along the batch
    constexpr int PX = TM / XG;         // rows per thread
    const int xb = tid % TK, xr0 = (tid / TK) * PX;
    constexpr int PY = TK * TN / 256;
    constexpr int YT = TK / PY;
    const int yrow = tid / YT, yb0 = (tid % YT) * PY;
    const int yk = kt0 + yrow;
    const int64_t ybase = (yk < n) ? (int64_t)crow[yk] * B : -1;
    const int64_t xbase = (nrow0 + xr0) * B;

    float xv[PX], yv[PY];
    auto load = [&](int64_t b0) {
        const int64_t b = b0 + xb;
#pragma unroll
        for (int q = 0; q < PX; ++q)
            xv[q] = (b < bend) ? lfdm_at(nflows, nmars, xbase + (int64_t)q * B + b, modified) : kNegInf;
#pragma unroll
        for (int q = 0; q < PY; ++q) {
            const int64_t b = b0 + yb0 + q;
            yv[q] = (ybase >= 0 && b < bend) ? emars[ybase + b] : kNegInf;
        }
    };

    floatx16 acc = {0};
    load(bstart);
    for (int64_t b0 = bstart; b0 < bend; b0 += TK) {
        float lm = kNegInf;
#pragma unroll
        for (int q = 0; q < PX; ++q) lm = fmaxf(lm, xv[q]);
        xred[tid / TK][xb] = lm;
        __syncthreads();
        if (tid < TK) {
            float t = xred[0][tid];
#pragma unroll
            for (int g = 1; g < XG; ++g) t = fmaxf(t, xred[g][tid]);
            sh[tid] = t;
        }
        __syncthreads();
        {
            const float s = sh[xb];
#pragma unroll
            for (int q = 0; q < PX; ++q)
                As[xr0 + q][xb] = MF<T>::cvt((s == kNegInf || xv[q] == kNegInf) ? 0.0f : __expf(xv[q] - s));
        }
#pragma unroll
        for (int q = 0; q < PY; ++q) {
            const float s = sh[yb0 + q];
            Bs[yrow][yb0 + q] = MF<T>::cvt((s == kNegInf || yv[q] == kNegInf) ? 0.0f : __expf(yv[q] + s));
        }
        __syncthreads();
        if (b0 + TK < bend) load(b0 + TK);
        acc = stage_mma<T, WM, WN, LD, LD>(As, Bs, wm, wn, lane, acc);
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
    // Split the batch only when the tile grid alone cannot fill the chip (<= 4 ways); split partial
    // sums are added atomically.
    int64_t splits = (512 + tiles - 1) / tiles;
    const int64_t max_splits = (B + TK - 1) / TK;
    if (splits > max_splits) splits = max_splits;
    if (splits > 4) splits = 4;
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
