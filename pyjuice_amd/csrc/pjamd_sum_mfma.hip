// Sum layers on MFMA (CDNA4 32x32 matrix cores): forward log-sum-exp, element flows and
// parameter flows for block sizes >= 32.  Operands are BF16, FP16 or FP32 (v_mfma_f32_32x32x16_bf16,
// v_mfma_f32_32x32x16_f16, v_mfma_f32_32x32x2_f32); accumulation is always FP32.
//
// Every kernel is one workgroup of 4 wavefronts computing a TM x TN tile (each wavefront one
// 32 x 32 MFMA accumulator).  Operand tiles are staged through LDS with K (the reduction index)
// contiguous, TK = 128 reduction steps per stage, and the next stage is prefetched into registers
// while the matrix cores work on the current one.  Deep stages matter because most layers of a
// tree-shaped circuit are narrow (a few node blocks): there the kernel time is the number of
// dependent memory round trips, not bandwidth.  Range is controlled by max subtraction:
//   forward   : exp(e - colmax_over_children)          (values in [0, 1])
//   ele-flows : exp(log(f / m) - colmax_over_parents)  (values in [0, 1])
//   par-flows : exp(log(f / m) - s_b) * exp(e + s_b),  s_b = max over the tile's parent rows
//
// Grids are one-dimensional over logical tiles.  The hardware hands workgroup i to XCD i % 8;
// xcd_tile() gives each XCD a contiguous range of logical tiles so that the tiles of one node
// block (which share parameters and child rows) meet in the same L2.
#include "pjamd_common.h"

using namespace pjamd;

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

constexpr int TK = 128;
constexpr int kXcds = 8;

template <typename T>
struct MF;

template <>
struct MF<__bf16> {
    static constexpr int KPL = 8;  // reduction values per lane per instruction (instruction K = 16)
    static constexpr int PAD = 8;  // LDS row stride 272 B: 16-byte aligned, 4-bank shift per row
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
    static constexpr int PAD = 1;  // LDS row stride 129 words: lanes of a column read distinct banks
    static __device__ __forceinline__ floatx16 mma(const float* a, const float* b, floatx16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(*a, *b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ float cvt(float x) { return x; }
};

// Row of accumulator register r for this lane inside a 32 x 32 tile (the column is lane & 31).
__device__ __forceinline__ int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

template <typename T, int LD>
__device__ __forceinline__ floatx16 stage_mma(const T (*As)[LD], const T (*Bs)[LD], int wm, int wn, int lane,
                                              floatx16 acc) {
#pragma unroll
    for (int kk = 0; kk < TK; kk += 2 * MF<T>::KPL) {
        acc = MF<T>::mma(&As[wm * 32 + (lane & 31)][kk + MF<T>::KPL * (lane >> 5)],
                         &Bs[wn * 32 + (lane & 31)][kk + MF<T>::KPL * (lane >> 5)], acc);
    }
    return acc;
}

// Logical tile of workgroup `bid` among `nb`: XCD x = bid % 8 owns the contiguous logical range
// [start(x), start(x) + count(x)), where the first nb % 8 XCDs own one tile more.  Bijective.
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
    const int per = nb / kXcds, rem = nb % kXcds;
    const int x = bid % kXcds, s = bid / kXcds;
    return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + s;
}

__device__ __forceinline__ float lfdm_at(const float* nflows, const float* nmars, int64_t o, bool modified) {
    if (modified) return nflows[o];
    const float m = nmars[o];
    return (m == kNegInf) ? kNegInf : __logf(nflows[o]) - m;
}

constexpr int KCAP = 1024;  // child / parent index entries staged in LDS per chunk

// ---------------------------------------------------------------------------------------------
// Forward.  Tile = (node block j, TM parent rows) x TN batch columns; logical tile order is
// m-tile fastest, then batch tile, then node block.  The column max is a running max
// (accumulators are rescaled when it grows).
//
// RUN: the caller promises that every TK-aligned group of index entries of a row names
// consecutive element rows (cids[k0 + i] = cids[k0] + i) whose parameters are bs apart
// (pids[k0 + i] = pids[k0] + i bs), which is how the compiler lays out a child block of size
// >= TK.  A stage then needs two indices instead of TK, read straight from memory; otherwise
// the index rows are staged through LDS.
template <typename T, int TM, int TN, bool RUN>
__global__ __launch_bounds__(256) void sum_fwd_mfma(float* __restrict__ node_mars, const float* __restrict__ emars,
                                                    const float* __restrict__ params, const int* __restrict__ nids,
                                                    const int* __restrict__ cids, const int* __restrict__ pids,
                                                    const int* __restrict__ nch, int K, int64_t B, int mtiles,
                                                    int btiles) {
    constexpr int WM = TM / 32, WN = TN / 32;
    static_assert(WM * WN == 4, "four wavefronts per workgroup");
    constexpr int LD = TK + MF<T>::PAD;
    constexpr int G = 256 / TN, PB = TK / G;       // element tile: G row groups x TN columns
    constexpr int PA = TK * TM / 256;              // parameter tile: PA consecutive k per thread
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    __shared__ int sc[RUN ? 1 : KCAP], sp[RUN ? 1 : KCAP];
    __shared__ float red[G][TN], mrun[TN], msc[TN];

    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int m0 = (t % mtiles) * TM;
    const int j = t / (mtiles * btiles);
    const int64_t b0 = (int64_t)((t / mtiles) % btiles) * TN;
    const int bs = mtiles * TM;
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

    // One stage of the reduction over children [k0, k0 + TK) of the current index chunk.
    auto process = [&](int k0, int cn, auto&& load) {
        float lm = kNegInf;
#pragma unroll
        for (int q = 0; q < PB; ++q) lm = fmaxf(lm, ev[q]);
        red[bg][bcol] = lm;
        __syncthreads();
        if (tid < TN) {
            float v = red[0][tid];
#pragma unroll
            for (int g = 1; g < G; ++g) v = fmaxf(v, red[g][tid]);
            const float mo = mrun[tid];
            const float mn = fmaxf(mo, v);
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
        acc = stage_mma<T, LD>(As, Bs, wm, wn, lane, acc);
    };

    if (RUN) {
        auto load = [&](int k0) {
            const int64_t c0 = crow[k0], p0 = prow[k0];
            const float* ep = emars + (c0 + bk0) * B + (bok ? bb : 0);
            const float* pp = params + p0 + (int64_t)ak0 * bs + m0 + arow;
            if (k0 + TK <= n && bok) {   // full stage: no per-element predicates
#pragma unroll
                for (int q = 0; q < PB; ++q) ev[q] = ep[(int64_t)q * B];
#pragma unroll
                for (int q = 0; q < PA; ++q) av[q] = pp[(int64_t)q * bs];
            } else {
#pragma unroll
                for (int q = 0; q < PB; ++q) ev[q] = (k0 + bk0 + q < n && bok) ? ep[(int64_t)q * B] : kNegInf;
#pragma unroll
                for (int q = 0; q < PA; ++q) av[q] = (k0 + ak0 + q < n) ? pp[(int64_t)q * bs] : 0.0f;
            }
        };
        load(0);
        for (int k0 = 0; k0 < n; k0 += TK) process(k0, n, load);
    } else {
        // Index rows are staged by their padded width K, so the first chunk's loads do not wait
        // for the child count n.
        for (int cb = 0; cb < n; cb += KCAP) {
            const int cn = (n - cb < KCAP) ? n - cb : KCAP;
            const int ck = (K - cb < KCAP) ? K - cb : KCAP;
            __syncthreads();
            for (int u = tid; u < ck; u += 256) {
                sc[u] = crow[cb + u];
                sp[u] = prow[cb + u];
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
            for (int k0 = 0; k0 < cn; k0 += TK) process(k0, cn, load);
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
// Element flows.  Tile = (child block j, TM child rows) x TN batch columns; the reduction runs
// over the flattened (parent block, parent row) index kk -> (kk >> lbs, kk & (bs - 1)) with a
// running max of log(f / m).  Parameters of one child row are contiguous along the parent row,
// so the parameter tile is loaded with lanes along kk.
//
// FUSE (children scatter): instead of element flows, the epilogue writes the flows of the product
// nodes' children straight into node_flows.  A product passes its flow unchanged to every child;
// the caller guarantees that each child node has that product as its only parent and that the
// child node blocks have the product block's size.  ch_nodes[j][k] = 2 * (first node of child
// block k of the product block j) + 1 if that child's flows are to be stored as
// log(flow) - mar (and mod_children is set), 0 = padding.
template <typename T, int TM, int TN, bool ONE_BLOCK, bool FUSE>
__global__ __launch_bounds__(256) void sum_ele_mfma(float* __restrict__ eflows, const float* nflows,
                                                    const float* __restrict__ nmars, const float* __restrict__ emars,
                                                    const float* __restrict__ params, const int* __restrict__ chids,
                                                    const int* __restrict__ parids, const int* __restrict__ parpids,
                                                    const int* __restrict__ npar, int KP, int lbs, int64_t B,
                                                    int mtiles, int btiles, bool modified, bool accumulate,
                                                    float* child_flows, const int* __restrict__ ch_nodes, int Kc,
                                                    bool mod_children) {
    constexpr int WM = TM / 32, WN = TN / 32;
    static_assert(WM * WN == 4, "four wavefronts per workgroup");
    constexpr int LD = TK + MF<T>::PAD;
    constexpr int G = 256 / TN, PB = TK / G;
    constexpr int RA = 256 / TK, PA = TM / RA;     // parameter tile: RA rows per pass, PA passes
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    __shared__ int spar[ONE_BLOCK ? 1 : KCAP], spp[ONE_BLOCK ? 1 : KCAP];
    __shared__ float red[G][TN], mrun[TN], msc[TN];

    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int m0 = (t % mtiles) * TM;
    const int j = t / (mtiles * btiles);
    const int64_t b0 = (int64_t)((t / mtiles) % btiles) * TN;
    const int bs = 1 << lbs;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w % WM, wn = w / WM;
    const int np = npar[j];
    const int* parrow = parids + (int64_t)j * KP;
    const int* pprow = parpids + (int64_t)j * KP;
    const int bcol = tid % TN, bg = tid / TN, bk0 = bg * PB;
    const int64_t bb = b0 + bcol;
    const bool bok = bb < B;
    const int ak = tid % TK, ar0 = tid / TK;
    const int mycol = wn * 32 + (lane & 31);

    if (tid < TN) mrun[tid] = kNegInf;
    floatx16 acc = {0};
    float lv[PB], av[PA];
    // The element marginals the epilogue needs do not depend on the reduction: load them now.
    const int64_t b = b0 + mycol;
    const int64_t obase = (int64_t)chids[j] + m0 + wm * 32;
    float epre[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) epre[r] = (b < B) ? emars[(obase + acc_row(r, lane)) * B + b] : kNegInf;
    // One stage of the reduction over the flattened parent index [k0, k0 + TK).
    auto process = [&](int k0, int kf, auto&& load) {
        float lm = kNegInf;
#pragma unroll
        for (int q = 0; q < PB; ++q) lm = fmaxf(lm, lv[q]);
        red[bg][bcol] = lm;
        __syncthreads();
        if (tid < TN) {
            float v = red[0][tid];
#pragma unroll
            for (int g = 1; g < G; ++g) v = fmaxf(v, red[g][tid]);
            const float mo = mrun[tid];
            const float mn = fmaxf(mo, v);
            msc[tid] = (mn == kNegInf) ? 1.0f : __expf(mo - mn);
            mrun[tid] = mn;
        }
        __syncthreads();
        const float mn = mrun[bcol];
#pragma unroll
        for (int q = 0; q < PB; ++q) Bs[bcol][bk0 + q] = MF<T>::cvt(mn == kNegInf ? 0.0f : __expf(lv[q] - mn));
#pragma unroll
        for (int q = 0; q < PA; ++q) As[ar0 + q * RA][ak] = MF<T>::cvt(av[q]);
        const float s = msc[mycol];
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] *= s;
        __syncthreads();
        if (k0 + TK < kf) load(k0 + TK);
        acc = stage_mma<T, LD>(As, Bs, wm, wn, lane, acc);
    };

    if (ONE_BLOCK) {
        // bs >= TK (both powers of two): a stage lies inside one parent block, so its rows and
        // parameters are contiguous runs and need only that block's two indices.
        const int kf = np << lbs;
        auto load = [&](int k0) {
            const int kb = k0 >> lbs, k_in = k0 & (bs - 1);
            const int64_t rbase = ((int64_t)parrow[kb] + k_in + bk0) * B + bb;
#pragma unroll
            for (int q = 0; q < PB; ++q)
                lv[q] = bok ? lfdm_at(nflows, nmars, rbase + (int64_t)q * B, modified) : kNegInf;
            const int64_t base = (int64_t)pprow[kb] + (int64_t)(m0 + ar0) * bs + k_in + ak;
#pragma unroll
            for (int q = 0; q < PA; ++q) av[q] = params[base + (int64_t)q * RA * bs];
        };
        load(0);
        for (int k0 = 0; k0 < kf; k0 += TK) process(k0, kf, load);
    } else {
        for (int pb = 0; pb < np; pb += KCAP) {
            const int cnp = (np - pb < KCAP) ? np - pb : KCAP;
            const int ckp = (KP - pb < KCAP) ? KP - pb : KCAP;
            const int kf = cnp << lbs;
            __syncthreads();
            for (int u = tid; u < ckp; u += 256) {
                spar[u] = parrow[pb + u];
                spp[u] = pprow[pb + u];
            }
            __syncthreads();
            auto load = [&](int k0) {
#pragma unroll
                for (int q = 0; q < PB; ++q) {
                    const int kk = k0 + bk0 + q;
                    float v = kNegInf;
                    if (kk < kf && bok)
                        v = lfdm_at(nflows, nmars, ((int64_t)spar[kk >> lbs] + (kk & (bs - 1))) * B + bb, modified);
                    lv[q] = v;
                }
                const int kk = k0 + ak;
                if (kk < kf) {
                    const int64_t base = (int64_t)spp[kk >> lbs] + (int64_t)(m0 + ar0) * bs + (kk & (bs - 1));
#pragma unroll
                    for (int q = 0; q < PA; ++q) av[q] = params[base + (int64_t)q * RA * bs];
                } else {
#pragma unroll
                    for (int q = 0; q < PA; ++q) av[q] = 0.0f;
                }
            };
            load(0);
            for (int k0 = 0; k0 < kf; k0 += TK) process(k0, kf, load);
        }
    }
    if (b < B) {
        const float m = mrun[mycol];
        if (FUSE) {
            float v[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float e = epre[r];
                v[r] = (m == kNegInf || e == kNegInf) ? 0.0f : acc[r] * __expf(e + m);
            }
            const int* crow = ch_nodes + (int64_t)j * Kc;
            for (int k = 0; k < Kc; ++k) {
                const int code = crow[k];
                if (code == 0) continue;
                const int64_t cbase = (int64_t)(code >> 1) + m0 + wm * 32;
                if ((code & 1) && mod_children) {
                    float mv[16];
#pragma unroll
                    for (int r = 0; r < 16; ++r) mv[r] = nmars[(cbase + acc_row(r, lane)) * B + b];
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        child_flows[(cbase + acc_row(r, lane)) * B + b] =
                            (mv[r] == kNegInf) ? kNegInf : __logf(v[r]) - mv[r];
                } else {
#pragma unroll
                    for (int r = 0; r < 16; ++r) child_flows[(cbase + acc_row(r, lane)) * B + b] = v[r];
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t o = (obase + acc_row(r, lane)) * B + b;
                const float e = epre[r];
                float v = (m == kNegInf || e == kNegInf) ? 0.0f : acc[r] * __expf(e + m);
                if (accumulate) v += eflows[o];
                eflows[o] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Parameter flows of one tile = (node block j, TM parent rows from m0) x (TN children from kt0),
// reduced over the batch range [bstart, bend).  Both operand tiles are loaded with lanes along
// the batch.  Shared by the per-layer and the batched kernels below.
template <typename T, int TM, int TN, typename PF>
__device__ __forceinline__ void par_tile(PF* __restrict__ pflows, const float* __restrict__ nflows,
                                         const float* __restrict__ nmars, const float* __restrict__ emars,
                                         const float* __restrict__ params, const int* __restrict__ nids,
                                         const int* __restrict__ cids, const int* __restrict__ pids,
                                         const int* __restrict__ pfids, const int* __restrict__ nch, int K, int64_t B,
                                         int j, int m0, int kt0, int64_t bstart, int64_t bend, bool modified,
                                         int atomic_out, float sign) {
    constexpr int WM = TM / 32, WN = TN / 32;
    static_assert(WM * WN == 4, "four wavefronts per workgroup");
    constexpr int LD = TK + MF<T>::PAD;
    constexpr int RG = 256 / TK;        // rows per pass of a tile load (lanes run along the batch)
    constexpr int PX = TM / RG, PY = TN / RG;
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    __shared__ float xred[RG][TK];
    __shared__ float sh[TK];
    __shared__ int64_t sy[TN];          // element row offsets of the tile's children (-1: padding)

    const int n = nch[j];
    if (kt0 >= n) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w % WM, wn = w / WM;
    const int64_t nrow0 = (int64_t)nids[j] + m0;
    const int* crow = cids + (int64_t)j * K;

    const int xb = tid % TK, r0 = tid / TK;
    if (tid < TN) sy[tid] = (kt0 + tid < n) ? (int64_t)crow[kt0 + tid] * B : -1;
    __syncthreads();
    const int64_t xbase = (nrow0 + r0) * B;

    float xv[PX], yv[PY];
    auto load = [&](int64_t b0) {
        const int64_t b = b0 + xb;
        const bool ok = b < bend;
#pragma unroll
        for (int q = 0; q < PX; ++q)
            xv[q] = ok ? lfdm_at(nflows, nmars, xbase + (int64_t)q * RG * B + b, modified) : kNegInf;
#pragma unroll
        for (int q = 0; q < PY; ++q) {
            const int64_t yb = sy[r0 + q * RG];
            yv[q] = (ok && yb >= 0) ? emars[yb + b] : kNegInf;
        }
    };

    floatx16 acc = {0};
    load(bstart);
    for (int64_t b0 = bstart; b0 < bend; b0 += TK) {
        float lm = kNegInf;
#pragma unroll
        for (int q = 0; q < PX; ++q) lm = fmaxf(lm, xv[q]);
        xred[r0][xb] = lm;
        __syncthreads();
        if (tid < TK) {
            float v = xred[0][tid];
#pragma unroll
            for (int g = 1; g < RG; ++g) v = fmaxf(v, xred[g][tid]);
            sh[tid] = v;
        }
        __syncthreads();
        const float s = sh[xb];
#pragma unroll
        for (int q = 0; q < PX; ++q)
            As[r0 + q * RG][xb] = MF<T>::cvt((s == kNegInf || xv[q] == kNegInf) ? 0.0f : __expf(xv[q] - s));
#pragma unroll
        for (int q = 0; q < PY; ++q)
            Bs[r0 + q * RG][xb] = MF<T>::cvt((s == kNegInf || yv[q] == kNegInf) ? 0.0f : __expf(yv[q] + s));
        __syncthreads();
        if (b0 + TK < bend) load(b0 + TK);
        acc = stage_mma<T, LD>(As, Bs, wm, wn, lane, acc);
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

// Per-layer launch: logical tile order is m-tile fastest, then child tile, then batch split,
// then node block.
template <typename T, int TM, int TN, typename PF>
__global__ __launch_bounds__(256) void sum_par_mfma(PF* __restrict__ pflows, const float* __restrict__ nflows,
                                                    const float* __restrict__ nmars, const float* __restrict__ emars,
                                                    const float* __restrict__ params, const int* __restrict__ nids,
                                                    const int* __restrict__ cids, const int* __restrict__ pids,
                                                    const int* __restrict__ pfids, const int* __restrict__ nch, int K,
                                                    int64_t B, int64_t bchunk, int mtiles, int ktiles, int splits,
                                                    bool modified, int atomic_out, float sign) {
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int m0 = (t % mtiles) * TM;
    const int kt0 = ((t / mtiles) % ktiles) * TN;
    const int z = (t / (mtiles * ktiles)) % splits;
    const int j = t / (mtiles * ktiles * splits);
    const int64_t bstart = (int64_t)z * bchunk;
    const int64_t bend = (bstart + bchunk < B) ? bstart + bchunk : B;
    par_tile<T, TM, TN, PF>(pflows, nflows, nmars, emars, params, nids, cids, pids, pfids, nch, K, B, j, m0, kt0,
                            bstart, bend, modified, atomic_out, sign);
}

// Batched launch over the partitions of many sum layers (one descriptor row per partition, see
// pjamd_sum_backward_par_batched).  The tile order inside a partition is m-tile fastest, then
// child tile, then node block; each tile reduces over the whole batch.
template <typename T, int TM, int TN, typename PF>
__global__ __launch_bounds__(256) void sum_par_batched(PF* __restrict__ pflows, const float* __restrict__ nflows,
                                                       const float* __restrict__ nmars,
                                                       const float* __restrict__ emars,
                                                       const float* __restrict__ params, const int* __restrict__ nids,
                                                       const int* __restrict__ cids, const int* __restrict__ pids,
                                                       const int* __restrict__ pfids, const int* __restrict__ nch,
                                                       const int* __restrict__ desc,
                                                       const int* __restrict__ tile_desc, int64_t B, bool modified,
                                                       bool overwrite, float sign) {
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int* D = desc + (int64_t)tile_desc[t] * PJAMD_PAR_DESC_FIELDS;
    const int local = t - D[0];
    const int K = D[1], mtiles = D[2], ktiles = D[3];
    const int64_t bo = D[4], io = D[5];
    const int m0 = (local % mtiles) * TM;
    const int kt0 = ((local / mtiles) % ktiles) * TN;
    const int j = local / (mtiles * ktiles);
    const int mode = D[7] ? kPfAtomic : (overwrite ? kPfOverwrite : kPfAdd);
    par_tile<T, TM, TN, PF>(pflows, nflows, nmars, emars + (int64_t)D[6] * B, params, nids + bo, cids + io, pids + io,
                            pfids + io, nch + bo, K, B, j, m0, kt0, 0, B, modified, mode, sign);
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

bool grid_fits(int64_t tiles) { return tiles > 0 && tiles <= (int64_t)0x7fffffff / 256; }

int log2_exact(int v) {
    if (v < 1 || (v & (v - 1)) != 0) return -1;
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

template <typename T, int TM, int TN>
void launch_fwd_tile(bool run, unsigned g, float* node_mars, const float* emars, const float* params, const int* nids,
                     const int* cids, const int* pids, const int* nch, int K, int64_t B, int mtiles, int btiles,
                     hipStream_t stream) {
    if (run)
        sum_fwd_mfma<T, TM, TN, true><<<g, 256, 0, stream>>>(node_mars, emars, params, nids, cids, pids, nch, K, B,
                                                             mtiles, btiles);
    else
        sum_fwd_mfma<T, TM, TN, false><<<g, 256, 0, stream>>>(node_mars, emars, params, nids, cids, pids, nch, K, B,
                                                              mtiles, btiles);
}

template <typename T>
int launch_fwd(float* node_mars, const float* emars, const float* params, const int* nids, const int* cids,
               const int* pids, const int* nch, int n_blocks, int K, int bs, int64_t B, int child_run,
               hipStream_t stream) {
    const int TM = bs >= 64 ? 64 : 32, TN = bs >= 64 ? 64 : 128;
    const int mtiles = bs / TM;
    const int64_t btiles = cdiv(B, TN);
    const int64_t tiles = (int64_t)n_blocks * mtiles * btiles;
    PJAMD_REQUIRE(bs % TM == 0 && grid_fits(tiles));
    const bool run = child_run >= TK && child_run % TK == 0;
    if (bs >= 64)
        launch_fwd_tile<T, 64, 64>(run, (unsigned)tiles, node_mars, emars, params, nids, cids, pids, nch, K, B, mtiles,
                                   (int)btiles, stream);
    else
        launch_fwd_tile<T, 32, 128>(run, (unsigned)tiles, node_mars, emars, params, nids, cids, pids, nch, K, B, mtiles,
                                    (int)btiles, stream);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

template <typename T, int TM, int TN, bool FUSE>
void launch_ele_tile(bool one_block, unsigned g, float* eflows, const float* nflows, const float* nmars,
                     const float* emars, const float* params, const int* chids, const int* parids,
                     const int* parpids, const int* npar, int KP, int lbs, int64_t B, int mtiles, int bt, bool mod,
                     bool accum, float* child_flows, const int* ch_nodes, int Kc, bool mod_children,
                     hipStream_t stream) {
    if (one_block)
        sum_ele_mfma<T, TM, TN, true, FUSE><<<g, 256, 0, stream>>>(eflows, nflows, nmars, emars, params, chids, parids,
                                                                   parpids, npar, KP, lbs, B, mtiles, bt, mod, accum,
                                                                   child_flows, ch_nodes, Kc, mod_children);
    else
        sum_ele_mfma<T, TM, TN, false, FUSE><<<g, 256, 0, stream>>>(eflows, nflows, nmars, emars, params, chids,
                                                                    parids, parpids, npar, KP, lbs, B, mtiles, bt, mod,
                                                                    accum, child_flows, ch_nodes, Kc, mod_children);
}

template <typename T>
int launch_ele(float* eflows, const float* nflows, const float* nmars, const float* emars, const float* params,
               const int* chids, const int* parids, const int* parpids, const int* npar, int n, int KP, int cbs, int bs,
               int64_t B, bool mod, bool accum, float* child_flows, const int* ch_nodes, int Kc, bool mod_children,
               hipStream_t stream) {
    const int lbs = log2_exact(bs);
    const int TM = cbs >= 64 ? 64 : 32;
    const int TN = cbs >= 64 ? 64 : 128;
    const int mtiles = cbs / TM;
    const int64_t btiles = cdiv(B, TN);
    const int64_t tiles = (int64_t)n * mtiles * btiles;
    PJAMD_REQUIRE(lbs >= 0 && cbs % TM == 0 && grid_fits(tiles));
    const unsigned g = (unsigned)tiles;
    const int bt = (int)btiles;
    const bool ob = bs >= TK;
    const bool fuse = ch_nodes != nullptr;
    if (cbs >= 64) {
        if (fuse)
            launch_ele_tile<T, 64, 64, true>(ob, g, eflows, nflows, nmars, emars, params, chids, parids, parpids, npar,
                                             KP, lbs, B, mtiles, bt, mod, accum, child_flows, ch_nodes, Kc,
                                             mod_children, stream);
        else
            launch_ele_tile<T, 64, 64, false>(ob, g, eflows, nflows, nmars, emars, params, chids, parids, parpids, npar,
                                              KP, lbs, B, mtiles, bt, mod, accum, child_flows, ch_nodes, Kc,
                                              mod_children, stream);
    } else {
        if (fuse)
            launch_ele_tile<T, 32, 128, true>(ob, g, eflows, nflows, nmars, emars, params, chids, parids, parpids, npar,
                                              KP, lbs, B, mtiles, bt, mod, accum, child_flows, ch_nodes, Kc,
                                              mod_children, stream);
        else
            launch_ele_tile<T, 32, 128, false>(ob, g, eflows, nflows, nmars, emars, params, chids, parids, parpids,
                                               npar, KP, lbs, B, mtiles, bt, mod, accum, child_flows, ch_nodes, Kc,
                                               mod_children, stream);
    }
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

template <typename T, typename PF>
int launch_par(PF* pflows, const float* nflows, const float* nmars, const float* emars, const float* params,
               const int* nids, const int* cids, const int* pids, const int* pfids, const int* nch, int n_blocks, int K,
               int bs, int64_t B, bool mod, int pf_mode, float sign, hipStream_t stream) {
    const int TM = bs >= 64 ? 64 : 32, TN = bs >= 64 ? 64 : 128;
    const int mtiles = bs / TM, ktiles = (K + TN - 1) / TN;
    const int64_t tiles = (int64_t)n_blocks * mtiles * ktiles;
    // Split the batch only when the tile grid alone cannot fill the chip (<= 4 ways); split partial
    // sums are added atomically.
    int64_t splits = (512 + tiles - 1) / tiles;
    const int64_t max_splits = (B + TK - 1) / TK;
    if (splits > max_splits) splits = max_splits;
    if (splits > 4) splits = 4;
    if (splits < 1 || pf_mode == kPfOverwrite) splits = 1;   // a stored slot must have one writer
    int64_t bchunk = (B + splits - 1) / splits;
    bchunk = ((bchunk + TK - 1) / TK) * TK;
    splits = (B + bchunk - 1) / bchunk;
    const int atomic_out = (pf_mode == kPfAtomic || splits > 1) ? kPfAtomic : pf_mode;
    PJAMD_REQUIRE(bs % TM == 0 && grid_fits(tiles * splits));
    const unsigned grid = (unsigned)(tiles * splits);
    if (bs >= 64)
        sum_par_mfma<T, 64, 64, PF><<<grid, 256, 0, stream>>>(pflows, nflows, nmars, emars, params, nids, cids, pids,
                                                             pfids, nch, K, B, bchunk, mtiles, ktiles, (int)splits, mod,
                                                             atomic_out, sign);
    else
        sum_par_mfma<T, 32, 128, PF><<<grid, 256, 0, stream>>>(pflows, nflows, nmars, emars, params, nids, cids, pids,
                                                              pfids, nch, K, B, bchunk, mtiles, ktiles, (int)splits,
                                                              mod, atomic_out, sign);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

template <typename T, typename PF>
int launch_par_batched(PF* pflows, const float* nflows, const float* nmars, const float* emars, const float* params,
                       const int* nids, const int* cids, const int* pids, const int* pfids, const int* nch,
                       const int* desc, const int* tile_desc, int n_tiles, int tile_rows, int64_t B, bool mod,
                       bool overwrite, float sign, hipStream_t stream) {
    PJAMD_REQUIRE(grid_fits(n_tiles));
    if (tile_rows == 64)
        sum_par_batched<T, 64, 64, PF><<<(unsigned)n_tiles, 256, 0, stream>>>(
            pflows, nflows, nmars, emars, params, nids, cids, pids, pfids, nch, desc, tile_desc, B, mod, overwrite, sign);
    else
        sum_par_batched<T, 32, 128, PF><<<(unsigned)n_tiles, 256, 0, stream>>>(
            pflows, nflows, nmars, emars, params, nids, cids, pids, pfids, nch, desc, tile_desc, B, mod, overwrite, sign);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

}  // namespace

// precision: 0 = FP32 operands, 1 = BF16, 2 = FP16 (FP32 accumulation in every case).
int pjamd_sum_forward_mfma(float* node_mars, const float* element_mars, const float* params, const int* nids,
                           const int* cids, const int* pids, const int* nch, int n_blocks, int K, int bs, int64_t B,
                           int child_run, int precision, hipStream_t stream) {
    if (bs < 32) return PJAMD_EUNSUPPORTED;
    if (precision == 1)
        return launch_fwd<__bf16>(node_mars, element_mars, params, nids, cids, pids, nch, n_blocks, K, bs, B,
                                  child_run, stream);
    if (precision == 2)
        return launch_fwd<_Float16>(node_mars, element_mars, params, nids, cids, pids, nch, n_blocks, K, bs, B,
                                    child_run, stream);
    return launch_fwd<float>(node_mars, element_mars, params, nids, cids, pids, nch, n_blocks, K, bs, B, child_run,
                             stream);
}

int pjamd_sum_backward_ele_mfma(float* eflows, const float* nflows, const float* nmars, const float* emars,
                                const float* params, const int* chids, const int* parids, const int* parpids,
                                const int* npar, int n, int KP, int cbs, int bs, int64_t B, bool mod, bool accum,
                                int precision, hipStream_t stream) {
    if (cbs < 32) return PJAMD_EUNSUPPORTED;
    if (precision == 1)
        return launch_ele<__bf16>(eflows, nflows, nmars, emars, params, chids, parids, parpids, npar, n, KP, cbs, bs, B,
                                  mod, accum, nullptr, nullptr, 0, false, stream);
    if (precision == 2)
        return launch_ele<_Float16>(eflows, nflows, nmars, emars, params, chids, parids, parpids, npar, n, KP, cbs, bs,
                                    B, mod, accum, nullptr, nullptr, 0, false, stream);
    return launch_ele<float>(eflows, nflows, nmars, emars, params, chids, parids, parpids, npar, n, KP, cbs, bs, B,
                             mod, accum, nullptr, nullptr, 0, false, stream);
}

PJAMD_EXPORT int pjamd_sum_backward_ele_to_children(float* node_flows, const float* node_mars,
                                                    const float* element_mars, const float* params, const int* chids,
                                                    const int* parids, const int* parpids, const int* n_pars,
                                                    const int* ch_nodes, int Kc, int n_cblocks, int KP, int cbs,
                                                    int bs, int64_t B, int flags, int precision, hipStream_t stream) {
    PJAMD_REQUIRE(n_cblocks >= 0 && KP >= 1 && Kc >= 1 && bs >= 1 && B > 0 && precision >= 0 && precision <= 2 &&
                  ch_nodes != nullptr);
    if (n_cblocks == 0) return PJAMD_OK;
    if (cbs < 32) return PJAMD_EUNSUPPORTED;
    const bool mod = flags & 1, modc = (flags & 2) != 0;
    if (precision == 1)
        return launch_ele<__bf16>(nullptr, node_flows, node_mars, element_mars, params, chids, parids, parpids, n_pars,
                                  n_cblocks, KP, cbs, bs, B, mod, false, node_flows, ch_nodes, Kc, modc, stream);
    if (precision == 2)
        return launch_ele<_Float16>(nullptr, node_flows, node_mars, element_mars, params, chids, parids, parpids,
                                    n_pars, n_cblocks, KP, cbs, bs, B, mod, false, node_flows, ch_nodes, Kc, modc,
                                    stream);
    return launch_ele<float>(nullptr, node_flows, node_mars, element_mars, params, chids, parids, parpids, n_pars,
                             n_cblocks, KP, cbs, bs, B, mod, false, node_flows, ch_nodes, Kc, modc, stream);
}

// FP16 operands can overflow in the parameter-flow product (exp(e + s) is unbounded above), so
// precision 2 runs this contraction in BF16.
int pjamd_sum_backward_par_mfma(void* pflows, bool pf64, const float* nflows, const float* nmars, const float* emars,
                                const float* params, const int* nids, const int* cids, const int* pids, const int* pfids,
                                const int* nch, int n_blocks, int K, int bs, int64_t B, bool mod, int shared, float sign,
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

PJAMD_EXPORT int pjamd_sum_backward_par_batched(void* param_flows, const float* node_flows, const float* node_mars,
                                                const float* element_mars, const float* params, const int* nids,
                                                const int* cids, const int* pids, const int* pfids, const int* nch,
                                                const int* desc, const int* tile_desc, int n_tiles, int tile_rows,
                                                int64_t B, int flags, int precision, hipStream_t stream) {
    PJAMD_REQUIRE(n_tiles >= 0 && B > 0 && (tile_rows == 32 || tile_rows == 64) && precision >= 0 && precision <= 2);
    if (n_tiles == 0) return PJAMD_OK;
    const bool mod = flags & 1;
    const float sign = (flags & 4) ? -1.0f : 1.0f;
    const bool pf64 = (flags & 8) != 0;
    const bool overwrite = (flags & 16) != 0;
    if (precision == 0) {
        if (pf64)
            return launch_par_batched<float, double>((double*)param_flows, node_flows, node_mars, element_mars, params,
                                                     nids, cids, pids, pfids, nch, desc, tile_desc, n_tiles, tile_rows,
                                                     B, mod, overwrite, sign, stream);
        return launch_par_batched<float, float>((float*)param_flows, node_flows, node_mars, element_mars, params, nids,
                                                cids, pids, pfids, nch, desc, tile_desc, n_tiles, tile_rows, B, mod,
                                                overwrite, sign, stream);
    }
    if (pf64)
        return launch_par_batched<__bf16, double>((double*)param_flows, node_flows, node_mars, element_mars, params,
                                                  nids, cids, pids, pfids, nch, desc, tile_desc, n_tiles, tile_rows, B,
                                                  mod, overwrite, sign, stream);
    return launch_par_batched<__bf16, float>((float*)param_flows, node_flows, node_mars, element_mars, params, nids,
                                             cids, pids, pfids, nch, desc, tile_desc, n_tiles, tile_rows, B, mod,
                                             overwrite, sign, stream);
}
