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
#include "pjamd_mfma.h"

#include <mutex>

using namespace pjamd;

namespace {

constexpr int TK = 128;
constexpr int TKB = 64;     // reduction steps per stage of the 128 x 128 tile kernels
constexpr int kXcds = 8;

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

constexpr int KCAP = 1024;  // child / parent index entries staged in LDS per chunk
constexpr int kPreChildren = 6;  // child blocks per product block the FUSE = 2 kernels prefetch

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
//
// RUN = 2 (grids smaller than the chip): the loads of the first two stages are issued together,
// so a two-stage reduction waits for one memory round trip instead of two.
template <typename T, int TM, int TN, int RUN>
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
        const float ns = neg_shift2(mn);   // exp(ev - mn); 0 for -inf entries
#pragma unroll
        for (int q = 0; q < PB; ++q) Bs[bcol][bk0 + q] = MF<T>::cvt(op_exp2(fmaf(ev[q], kLog2e, ns)));
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
        auto load_into = [&](int k0, float* evp, float* avp) {
            const int64_t c0 = crow[k0], p0 = prow[k0];
            const float* ep = emars + (c0 + bk0) * B + (bok ? bb : 0);
            const float* pp = params + p0 + (int64_t)ak0 * bs + m0 + arow;
            if (k0 + TK <= n && bok) {   // full stage: no per-element predicates
#pragma unroll
                for (int q = 0; q < PB; ++q) evp[q] = ep[(int64_t)q * B];
#pragma unroll
                for (int q = 0; q < PA; ++q) avp[q] = pp[(int64_t)q * bs];
            } else {
#pragma unroll
                for (int q = 0; q < PB; ++q) evp[q] = (k0 + bk0 + q < n && bok) ? ep[(int64_t)q * B] : kNegInf;
#pragma unroll
                for (int q = 0; q < PA; ++q) avp[q] = (k0 + ak0 + q < n) ? pp[(int64_t)q * bs] : 0.0f;
            }
        };
        float ev2[RUN == 2 ? PB : 1], av2[RUN == 2 ? PA : 1];
        load_into(0, ev, av);
        if (RUN == 2 && TK < n) load_into(TK, ev2, av2);
        auto next = [&](int k1) {
            if (RUN == 2 && k1 == TK) {
#pragma unroll
                for (int q = 0; q < PB; ++q) ev[q] = ev2[q];
#pragma unroll
                for (int q = 0; q < PA; ++q) av[q] = av2[q];
            } else {
                load_into(k1, ev, av);
            }
        };
        for (int k0 = 0; k0 < n; k0 += TK) process(k0, n, next);
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
//
// FUSE = 2 (at most kPreChildren child blocks per product block; chosen for grids smaller than the
// chip, where the kernel time is its chain of dependent memory round trips): the child codes and
// the marginals of the marked children are fetched at kernel start, beside the element marginals,
// instead of after the reduction.
// Children scatter of the fused product backward (FUSE = 1): this lane's 16 product-row flows
// v[r] (rows acc_row(r, lane) from roff, column b) go to the same rows of each of the product
// block's Kc child blocks, as log(flow) - mar for a marked child when the children are modified.
// In chunks of kScatterChunk child blocks, the chunk's codes and then the marginals of its marked
// children are all loaded before its first store, so a chunk waits for two memory round trips
// (one when its codes come in `first`, loaded at kernel start) instead of two per child.
constexpr int kScatterChunk = 2;

__device__ __forceinline__ void scatter_children(float* child_flows, const float* __restrict__ nmars,
                                                 const int* __restrict__ crow, int Kc, int64_t roff, int lane,
                                                 int64_t b, int64_t B, bool mod_children, const float (&v)[16],
                                                 const int* first = nullptr) {
    for (int k0 = 0; k0 < Kc; k0 += kScatterChunk) {
        int code[kScatterChunk];
#pragma unroll
        for (int u = 0; u < kScatterChunk; ++u)
            code[u] = (k0 == 0 && first != nullptr) ? first[u] : (k0 + u < Kc ? crow[k0 + u] : 0);
        float mv[kScatterChunk][16];
#pragma unroll
        for (int u = 0; u < kScatterChunk; ++u) {
            if ((code[u] & 1) && mod_children) {
                const int64_t cbase = (int64_t)(code[u] >> 1) + roff;
#pragma unroll
                for (int r = 0; r < 16; ++r) mv[u][r] = nmars[(cbase + acc_row(r, lane)) * B + b];
            }
        }
#pragma unroll
        for (int u = 0; u < kScatterChunk; ++u) {
            if (code[u] == 0) continue;
            const int64_t cbase = (int64_t)(code[u] >> 1) + roff;
            if ((code[u] & 1) && mod_children) {
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    child_flows[(cbase + acc_row(r, lane)) * B + b] =
                        (mv[u][r] == kNegInf) ? kNegInf : __logf(v[r]) - mv[u][r];
            } else {
#pragma unroll
                for (int r = 0; r < 16; ++r) child_flows[(cbase + acc_row(r, lane)) * B + b] = v[r];
            }
        }
    }
}

template <typename T, int TM, int TN, bool ONE_BLOCK, int FUSE>
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
    int cpre[FUSE == 2 ? kPreChildren : 1];
    float mpre[FUSE == 2 ? kPreChildren : 1][16];
    auto prefetch_epilogue = [&]() {
#pragma unroll
        for (int r = 0; r < 16; ++r) epre[r] = (b < B) ? emars[(obase + acc_row(r, lane)) * B + b] : kNegInf;
        if (FUSE == 2) {
            const int* crow = ch_nodes + (int64_t)j * Kc;
            const int64_t bc = b < B ? b : B - 1;   // clamped column: unconditional loads
#pragma unroll
            for (int u = 0; u < kPreChildren; ++u) {
                cpre[u] = u < Kc ? crow[u] : 0;
                if ((cpre[u] & 1) && mod_children) {
                    const int64_t cb = (int64_t)(cpre[u] >> 1) + m0 + wm * 32;
#pragma unroll
                    for (int r = 0; r < 16; ++r) mpre[u][r] = nmars[(cb + acc_row(r, lane)) * B + bc];
                }
            }
        }
    };
    // FUSE = 2 issues the reduction's loads first (all of them when it has two stages) and the
    // epilogue's after them; the others fetch the epilogue's operands first.
    if (FUSE != 2) prefetch_epilogue();
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
        const float ns = neg_shift2(mn);   // exp(lv - mn); 0 for -inf entries
#pragma unroll
        for (int q = 0; q < PB; ++q) Bs[bcol][bk0 + q] = MF<T>::cvt(op_exp2(fmaf(lv[q], kLog2e, ns)));
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
        auto load_into = [&](int k0, float* lvp, float* avp) {
            const int kb = k0 >> lbs, k_in = k0 & (bs - 1);
            const int64_t rbase = ((int64_t)parrow[kb] + k_in + bk0) * B + bb;
#pragma unroll
            for (int q = 0; q < PB; ++q)
                lvp[q] = bok ? lfdm_at(nflows, nmars, rbase + (int64_t)q * B, modified) : kNegInf;
            const int64_t base = (int64_t)pprow[kb] + (int64_t)(m0 + ar0) * bs + k_in + ak;
#pragma unroll
            for (int q = 0; q < PA; ++q) avp[q] = params[base + (int64_t)q * RA * bs];
        };
        float lv2[FUSE == 2 ? PB : 1], av2[FUSE == 2 ? PA : 1];
        load_into(0, lv, av);
        if (FUSE == 2) {
            if (TK < kf) load_into(TK, lv2, av2);
            prefetch_epilogue();
        }
        auto next = [&](int k1) {
            if (FUSE == 2 && k1 == TK) {
#pragma unroll
                for (int q = 0; q < PB; ++q) lv[q] = lv2[q];
#pragma unroll
                for (int q = 0; q < PA; ++q) av[q] = av2[q];
            } else {
                load_into(k1, lv, av);
            }
        };
        for (int k0 = 0; k0 < kf; k0 += TK) process(k0, kf, next);
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
        if (FUSE == 2) {
            float v[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float e = epre[r];
                v[r] = (m == kNegInf || e == kNegInf) ? 0.0f : acc[r] * __expf(e + m);
            }
#pragma unroll
            for (int u = 0; u < kPreChildren; ++u) {
                if (cpre[u] == 0) continue;
                const int64_t cbase = (int64_t)(cpre[u] >> 1) + m0 + wm * 32;
                if ((cpre[u] & 1) && mod_children) {
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        child_flows[(cbase + acc_row(r, lane)) * B + b] =
                            (mpre[u][r] == kNegInf) ? kNegInf : __logf(v[r]) - mpre[u][r];
                } else {
#pragma unroll
                    for (int r = 0; r < 16; ++r) child_flows[(cbase + acc_row(r, lane)) * B + b] = v[r];
                }
            }
        } else if (FUSE) {
            float v[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float e = epre[r];
                v[r] = (m == kNegInf || e == kNegInf) ? 0.0f : acc[r] * __expf(e + m);
            }
            scatter_children(child_flows, nmars, ch_nodes + (int64_t)j * Kc, Kc, m0 + wm * 32, lane, b, B, mod_children,
                             v);
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
// 16-byte-load forms of the two hot reductions (BF16 / FP16 operands, 64 x 64 tiles, TK = 128).
// Each thread moves a stage's operands with 16 float4 loads instead of 64 dword loads, so a
// wavefront stays well inside its 63 outstanding vector loads and a stage costs one memory round
// trip.  Requirements (checked by the launchers): B % 4 == 0; every parameter index of the
// partition a multiple of 4 (16-byte-aligned parameter rows); forward: the child-run layout with
// runs >= TK; element flows: parent blocks of >= TK rows (a stage inside one parent block).
//
// Element / flow rows (TK rows x 64 columns): thread = (row group g = tid / 16 of 8 rows,
// column quad q = tid % 16); its 8 float4 give 4 columns x 8 consecutive rows, which are one
// 16-byte LDS store per column of the [column][k] B tile.
constexpr int kVecRows = 8;                 // rows per thread of a row tile
constexpr int kVecGroups = 256 / 16;        // row groups (16 column quads per 64 columns)

template <typename T>
struct Pack8 {
    static __device__ __forceinline__ void store(T* dst, const float* v) {
        typedef T T8 __attribute__((ext_vector_type(8)));
        T8 x;
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = MF<T>::cvt(v[i]);
        *reinterpret_cast<T8*>(dst) = x;
    }
};

// Column max of a 128 x 64 stage held as v[8 rows][4 columns] per thread, folded into the running
// max; returns this thread's 4 column maxima (new running max).
__device__ __forceinline__ void vec_stage_max(const float4 (&v)[kVecRows], float (&red)[kVecGroups][64],
                                              float* mrun, float* msc, int g, int q, float (&mn)[4]) {
    float m[4] = {kNegInf, kNegInf, kNegInf, kNegInf};
#pragma unroll
    for (int r = 0; r < kVecRows; ++r) {
        m[0] = fmaxf(m[0], v[r].x);
        m[1] = fmaxf(m[1], v[r].y);
        m[2] = fmaxf(m[2], v[r].z);
        m[3] = fmaxf(m[3], v[r].w);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) red[g][4 * q + c] = m[c];
    __syncthreads();
    const int tid = threadIdx.x;
    if (tid < 64) {
        float x = red[0][tid];
#pragma unroll
        for (int i = 1; i < kVecGroups; ++i) x = fmaxf(x, red[i][tid]);
        const float mo = mrun[tid];
        const float mnew = fmaxf(mo, x);
        msc[tid] = (mnew == kNegInf) ? 1.0f : __expf(mo - mnew);
        mrun[tid] = mnew;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) mn[c] = mrun[4 * q + c];
}

// B tile from a stage: Bs[4q + c][8g .. 8g + 7] = exp(v[r][c] - mn[c]).
template <typename T, int LD>
__device__ __forceinline__ void vec_stage_b(const float4 (&v)[kVecRows], const float (&mn)[4], T (*Bs)[LD], int g,
                                            int q) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float e[8];
        const float ns = neg_shift2(mn[c]);   // exp(x - max); 0 for -inf entries
#pragma unroll
        for (int r = 0; r < kVecRows; ++r) {
            const float x = c == 0 ? v[r].x : c == 1 ? v[r].y : c == 2 ? v[r].z : v[r].w;
            e[r] = op_exp2(fmaf(x, kLog2e, ns));
        }
        Pack8<T>::store(&Bs[4 * q + c][8 * g], e);
    }
}

// Forward with 16-byte loads (RUN layout).  PRE: the first two stages' loads are issued together
// (grids smaller than the chip).
//
// KP > 0 (product layer fused in): the element rows are not read from emars but computed, as the
// product layer below would, from the node rows of the products' children: stage s of node block j
// covers element rows inside one product block, and pch[(j * nstages + s) * KP + u] is the first
// node row of the u-th child block of that product block, offset to the stage's first element
// row (padding points into the dummy node rows, which hold 0).  The workgroups of the first
// row tile (m0 = 0) also store the element rows into ewrite (the slab the backward reads).
template <typename T, bool PRE, int KP>
__global__ __launch_bounds__(256) void sum_fwd_vec(float* __restrict__ node_mars, const float* __restrict__ emars,
                                                   const float* __restrict__ params, const int* __restrict__ nids,
                                                   const int* __restrict__ cids, const int* __restrict__ pids,
                                                   const int* __restrict__ nch, int K, int64_t B, int mtiles,
                                                   int btiles, const float* __restrict__ nsrc,
                                                   const int* __restrict__ pch, int nstages, float* ewrite) {
    constexpr int TM = 64, TN = 64, LD = TK + MF<T>::PAD;
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    __shared__ float red[kVecGroups][64], mrun[TN], msc[TN];

    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int m0 = (t % mtiles) * TM;
    const int j = t / (mtiles * btiles);
    const int64_t b0 = (int64_t)((t / mtiles) % btiles) * TN;
    const int bs = mtiles * TM;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w % 2, wn = w / 2;
    const int n = nch[j];
    const int* crow = cids + (int64_t)j * K;
    const int* prow = pids + (int64_t)j * K;
    const int g = tid / 16, q = tid % 16;           // element rows: row group, column quad
    const int64_t bq = b0 + 4 * q;
    const bool bok = bq < B;                        // B % 4 == 0: a quad is all in or all out
    const int64_t bqc = bok ? bq : B - 4;
    const int iq = tid % 16, kg = tid / 16;         // parameters: row quad, child group of 8
    const int mycol = wn * 32 + (lane & 31);
    if (tid < TN) mrun[tid] = kNegInf;
    floatx16 acc = {0};

    float4 ev[kVecRows], av[kVecRows], ev2[PRE ? kVecRows : 1], av2[PRE ? kVecRows : 1];
    int64_t ce = 0, ce2 = 0;     // first element row of the stage held in ev / ev2
    auto load = [&](int k0, float4* e, float4* a, int64_t& ec) {
        const int64_t c0 = crow[k0], p0 = prow[k0];
        ec = c0;
        if constexpr (KP > 0) {
            const int* pb = pch + ((int64_t)j * nstages + k0 / TK) * KP;
            int64_t cb[KP];
#pragma unroll
            for (int u = 0; u < KP; ++u) cb[u] = pb[u];
#pragma unroll
            for (int r = 0; r < kVecRows; ++r) {
                const int k = 8 * g + r;
                const int kc = (k0 + k < n) ? k : 0;
                float4 v = *reinterpret_cast<const float4*>(nsrc + (cb[0] + kc) * B + bqc);
#pragma unroll
                for (int u = 1; u < KP; ++u) {
                    const float4 x = *reinterpret_cast<const float4*>(nsrc + (cb[u] + kc) * B + bqc);
                    v.x += x.x; v.y += x.y; v.z += x.z; v.w += x.w;
                }
                e[r] = v;
            }
        } else {
#pragma unroll
            for (int r = 0; r < kVecRows; ++r) {
                const int k = 8 * g + r;
                const int kc = (k0 + k < n) ? k : 0;
                e[r] = *reinterpret_cast<const float4*>(emars + (c0 + kc) * B + bqc);
                if (!(k0 + k < n && bok)) e[r] = make_float4(kNegInf, kNegInf, kNegInf, kNegInf);
            }
        }
#pragma unroll
        for (int r = 0; r < kVecRows; ++r) {
            const int k = 8 * kg + r;
            const int kc = (k0 + k < n) ? k : 0;
            a[r] = *reinterpret_cast<const float4*>(params + p0 + (int64_t)kc * bs + m0 + 4 * iq);
            if (!(k0 + k < n)) a[r] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    };
    if (n > 0) load(0, ev, av, ce);
    if (PRE && TK < n) load(TK, ev2, av2, ce2);
    for (int k0 = 0; k0 < n; k0 += TK) {
        if constexpr (KP > 0) {
            // the product rows of this stage: stored once (first row tile), masked for the reduction
#pragma unroll
            for (int r = 0; r < kVecRows; ++r) {
                const int k = 8 * g + r;
                const bool ok = k0 + k < n && bok;
                if (ok && m0 == 0) *reinterpret_cast<float4*>(ewrite + (ce + k) * B + bq) = ev[r];
                if (!ok) ev[r] = make_float4(kNegInf, kNegInf, kNegInf, kNegInf);
            }
        }
        float mn[4];
        vec_stage_max(ev, red, mrun, msc, g, q, mn);
        vec_stage_b<T, LD>(ev, mn, Bs, g, q);
        // As[4 iq + c][8 kg .. 8 kg + 7]
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float x[8];
#pragma unroll
            for (int r = 0; r < kVecRows; ++r)
                x[r] = c == 0 ? av[r].x : c == 1 ? av[r].y : c == 2 ? av[r].z : av[r].w;
            Pack8<T>::store(&As[4 * iq + c][8 * kg], x);
        }
        const float s = msc[mycol];
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] *= s;
        __syncthreads();
        if (k0 + TK < n) {
            if (PRE && k0 == 0) {
#pragma unroll
                for (int r = 0; r < kVecRows; ++r) {
                    ev[r] = ev2[r];
                    av[r] = av2[r];
                }
                ce = ce2;
            } else {
                load(k0 + TK, ev, av, ce);
            }
        }
        acc = stage_mma<T, LD>(As, Bs, wm, wn, lane, acc);
        __syncthreads();
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

// Forward with 128 x 128 tiles (16-bit operands, child runs >= TK, aligned parameter rows,
// B % 4 == 0, block sizes >= 128), for wide layers.  A 64 x 64 tile moves (64 + 64) K floats
// from L2 for 4096 outputs; 128 x 128 tiles halve that per output (the limit of the 64 x 64 form
// on wide layers: a CU keeps a bounded number of L2 requests in flight).  Four wavefronts, each a
// 64 x 64 quadrant of four 32 x 32 accumulators; stages of TKB = 64 children.  Element rows:
// thread (g = tid / 32, q = tid % 32) loads children 8 g .. 8 g + 7 at columns 4 q .. 4 q + 3;
// parameters: thread (iq = tid % 32, kg = tid / 32) loads node rows 4 iq .. 4 iq + 3 of children
// 8 kg .. 8 kg + 7.  The stage's column max goes through LDS once; every thread then folds it
// into the running maxima it needs (its four load columns and its two accumulator columns), so a
// stage has two barriers.
template <typename T>
__global__ __launch_bounds__(256) void sum_fwd_big(float* __restrict__ node_mars, const float* __restrict__ emars,
                                                   const float* __restrict__ params, const int* __restrict__ nids,
                                                   const int* __restrict__ cids, const int* __restrict__ pids,
                                                   const int* __restrict__ nch, int K, int64_t B, int mtiles,
                                                   int btiles) {
    constexpr int TM = 128, TN = 128, LD = TKB + MF<T>::PAD;
    __shared__ __attribute__((aligned(16))) T As[TM][LD];     // parameters [node row][child]
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];     // exp(e - max) [column][child]
    __shared__ __attribute__((aligned(16))) float red[4][TN];  // per-wavefront stage column max

    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int m0 = (t % mtiles) * TM;
    const int j = t / (mtiles * btiles);
    const int64_t b0 = (int64_t)((t / mtiles) % btiles) * TN;
    const int bs = mtiles * TM;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w & 1, wn = w >> 1;
    const int n = nch[j];
    const int* crow = cids + (int64_t)j * K;
    const int* prow = pids + (int64_t)j * K;
    const int g = tid >> 5, q = tid & 31;
    const int64_t bq = b0 + 4 * q;
    const bool bok = bq < B;                        // B % 4 == 0: a quad is all in or all out
    const int64_t bqc = bok ? bq : B - 4;
    const int iq = tid & 31, kg = tid >> 5;
    const float4 ninf = make_float4(kNegInf, kNegInf, kNegInf, kNegInf);
    floatx16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = floatx16{0};
    float mq[4] = {kNegInf, kNegInf, kNegInf, kNegInf};   // running max of columns 4 q + c
    float ma[2] = {kNegInf, kNegInf};                     // running max of accumulator columns

    float4 ev[8], av[8];
    auto load = [&](int k0) {
        // child run layout: the stage's children are consecutive element rows with parameters bs apart
        const int64_t c0 = crow[k0], p0 = prow[k0];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int k = 8 * g + r;
            const int kc = (k0 + k < n) ? k : 0;
            ev[r] = *reinterpret_cast<const float4*>(emars + (c0 + kc) * B + bqc);
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int k = 8 * kg + r;
            const int kc = (k0 + k < n) ? k : 0;
            av[r] = *reinterpret_cast<const float4*>(params + p0 + (int64_t)kc * bs + m0 + 4 * iq);
        }
    };
    if (n > 0) load(0);
    for (int k0 = 0; k0 < n; k0 += TKB) {
        // padding children and columns past B are -inf (element rows) and 0 (parameters)
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            if (!(k0 + 8 * g + r < n && bok)) ev[r] = ninf;
            if (!(k0 + 8 * kg + r < n)) av[r] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        float m4[4] = {kNegInf, kNegInf, kNegInf, kNegInf};
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            m4[0] = fmaxf(m4[0], ev[r].x);
            m4[1] = fmaxf(m4[1], ev[r].y);
            m4[2] = fmaxf(m4[2], ev[r].z);
            m4[3] = fmaxf(m4[3], ev[r].w);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) m4[c] = fmaxf(m4[c], __shfl_xor(m4[c], 32, 64));
        if (lane < 32) *reinterpret_cast<float4*>(&red[w][4 * q]) = make_float4(m4[0], m4[1], m4[2], m4[3]);
        __syncthreads();
        {
            float4 x = *reinterpret_cast<const float4*>(&red[0][4 * q]);
#pragma unroll
            for (int v = 1; v < 4; ++v) {
                const float4 y = *reinterpret_cast<const float4*>(&red[v][4 * q]);
                x = make_float4(fmaxf(x.x, y.x), fmaxf(x.y, y.y), fmaxf(x.z, y.z), fmaxf(x.w, y.w));
            }
            mq[0] = fmaxf(mq[0], x.x);
            mq[1] = fmaxf(mq[1], x.y);
            mq[2] = fmaxf(mq[2], x.z);
            mq[3] = fmaxf(mq[3], x.w);
        }
#pragma unroll
        for (int sj = 0; sj < 2; ++sj) {
            const int col = wn * 64 + sj * 32 + (lane & 31);
            const float x = fmaxf(fmaxf(red[0][col], red[1][col]), fmaxf(red[2][col], red[3][col]));
            const float mo = ma[sj], mn = fmaxf(mo, x);
            const float sc = (mn == kNegInf) ? 1.0f : __expf(mo - mn);
            ma[sj] = mn;
#pragma unroll
            for (int si = 0; si < 2; ++si)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[si][sj][r] *= sc;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float e[8];
            const float ns = neg_shift2(mq[c]);   // exp(e - max); 0 for -inf entries
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const float x = c == 0 ? ev[r].x : c == 1 ? ev[r].y : c == 2 ? ev[r].z : ev[r].w;
                e[r] = op_exp2(fmaf(x, kLog2e, ns));
            }
            Pack8<T>::store(&Bs[4 * q + c][8 * g], e);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float x[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) x[r] = c == 0 ? av[r].x : c == 1 ? av[r].y : c == 2 ? av[r].z : av[r].w;
            Pack8<T>::store(&As[4 * iq + c][8 * kg], x);
        }
        __syncthreads();
        if (k0 + TKB < n) load(k0 + TKB);
#pragma unroll
        for (int kk = 0; kk < TKB; kk += 2 * MF<T>::KPL) {
            const int kc = kk + MF<T>::KPL * (lane >> 5);
            const T* a0 = &As[wm * 64 + (lane & 31)][kc];
            const T* a1 = &As[wm * 64 + 32 + (lane & 31)][kc];
            const T* c0 = &Bs[wn * 64 + (lane & 31)][kc];
            const T* c1 = &Bs[wn * 64 + 32 + (lane & 31)][kc];
            acc[0][0] = MF<T>::mma(a0, c0, acc[0][0]);
            acc[0][1] = MF<T>::mma(a0, c1, acc[0][1]);
            acc[1][0] = MF<T>::mma(a1, c0, acc[1][0]);
            acc[1][1] = MF<T>::mma(a1, c1, acc[1][1]);
        }
    }
    const int64_t base = (int64_t)nids[j] + m0 + wm * 64;
#pragma unroll
    for (int sj = 0; sj < 2; ++sj) {
        const int64_t b = b0 + wn * 64 + sj * 32 + (lane & 31);
        if (b >= B) continue;
        const float m = ma[sj];
#pragma unroll
        for (int si = 0; si < 2; ++si)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float v = acc[si][sj][r];
                node_mars[(base + si * 32 + acc_row(r, lane)) * B + b] =
                    (v > 0.0f && m != kNegInf) ? __logf(v) + m : kNegInf;
            }
    }
}

// Element flows with 16-byte loads (parent blocks of >= TK rows).  FUSE as sum_ele_mfma (0, 1, or
// 2: child codes and the marked children's marginals fetched at kernel start, used with PRE);
// PRE as sum_fwd_vec.
template <typename T, int FUSE, bool PRE>
__global__ __launch_bounds__(256) void sum_ele_vec(float* __restrict__ eflows, const float* nflows,
                                                   const float* __restrict__ nmars, const float* __restrict__ emars,
                                                   const float* __restrict__ params, const int* __restrict__ chids,
                                                   const int* __restrict__ parids, const int* __restrict__ parpids,
                                                   const int* __restrict__ npar, int KP, int lbs, int64_t B,
                                                   int mtiles, int btiles, bool modified, bool accumulate,
                                                   float* child_flows, const int* __restrict__ ch_nodes, int Kc,
                                                   bool mod_children) {
    constexpr int TM = 64, TN = 64, LD = TK + MF<T>::PAD;
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    __shared__ float red[kVecGroups][64], mrun[TN], msc[TN];

    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int m0 = (t % mtiles) * TM;
    const int j = t / (mtiles * btiles);
    const int64_t b0 = (int64_t)((t / mtiles) % btiles) * TN;
    const int bs = 1 << lbs;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w % 2, wn = w / 2;
    const int np = npar[j];
    const int* parrow = parids + (int64_t)j * KP;
    const int* pprow = parpids + (int64_t)j * KP;
    const int g = tid / 16, q = tid % 16;            // parent rows: row group, column quad
    const int64_t bq = b0 + 4 * q;
    const bool bok = bq < B;
    const int64_t bqc = bok ? bq : B - 4;
    const int kq = tid % 32, ig = tid / 32;          // parameters: parent quad, child row group
    const int mycol = wn * 32 + (lane & 31);
    if (tid < TN) mrun[tid] = kNegInf;
    floatx16 acc = {0};

    const int kf = np << lbs;
    float4 lv[kVecRows], av[kVecRows], lv2[PRE ? kVecRows : 1], av2[PRE ? kVecRows : 1];
    auto load = [&](int k0, float4* l, float4* a) {
        const int kb = k0 >> lbs, k_in = k0 & (bs - 1);
        const int64_t rbase = (int64_t)parrow[kb] + k_in + 8 * g;
#pragma unroll
        for (int r = 0; r < kVecRows; ++r) {
            const int64_t o = (rbase + r) * B + bqc;
            float4 f = *reinterpret_cast<const float4*>(nflows + o);
            if (!modified) {
                const float4 m = *reinterpret_cast<const float4*>(nmars + o);
                f.x = m.x == kNegInf ? kNegInf : __logf(f.x) - m.x;
                f.y = m.y == kNegInf ? kNegInf : __logf(f.y) - m.y;
                f.z = m.z == kNegInf ? kNegInf : __logf(f.z) - m.z;
                f.w = m.w == kNegInf ? kNegInf : __logf(f.w) - m.w;
            }
            l[r] = bok ? f : make_float4(kNegInf, kNegInf, kNegInf, kNegInf);
        }
        const float* pp = params + (int64_t)pprow[kb] + (int64_t)(m0 + ig) * bs + k_in + 4 * kq;
#pragma unroll
        for (int r = 0; r < kVecRows; ++r) a[r] = *reinterpret_cast<const float4*>(pp + (int64_t)(8 * r) * bs);
    };
    if (kf > 0) load(0, lv, av);
    if (PRE && TK < kf) load(TK, lv2, av2);
    // element marginals for the epilogue
    const int64_t b = b0 + mycol;
    const int64_t obase = (int64_t)chids[j] + m0 + wm * 32;
    float epre[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) epre[r] = (b < B) ? emars[(obase + acc_row(r, lane)) * B + b] : kNegInf;
    int cpre[FUSE == 2 ? kPreChildren : 1];
    float mpre[FUSE == 2 ? kPreChildren : 1][16];
    // FUSE = 1: the first chunk of child codes (uniform), fetched beside the reduction's first loads
    int cfirst[kScatterChunk];
    if (FUSE == 1) {
        const int* crow = ch_nodes + (int64_t)j * Kc;
#pragma unroll
        for (int u = 0; u < kScatterChunk; ++u) cfirst[u] = u < Kc ? crow[u] : 0;
    }
    if (FUSE == 2) {
        const int* crow = ch_nodes + (int64_t)j * Kc;
        const int64_t bc = b < B ? b : B - 1;   // clamped column: unconditional loads
#pragma unroll
        for (int u = 0; u < kPreChildren; ++u) {
            cpre[u] = u < Kc ? crow[u] : 0;
            if ((cpre[u] & 1) && mod_children) {
                const int64_t cb = (int64_t)(cpre[u] >> 1) + m0 + wm * 32;
#pragma unroll
                for (int r = 0; r < 16; ++r) mpre[u][r] = nmars[(cb + acc_row(r, lane)) * B + bc];
            }
        }
    }
    for (int k0 = 0; k0 < kf; k0 += TK) {
        float mn[4];
        vec_stage_max(lv, red, mrun, msc, g, q, mn);
        vec_stage_b<T, LD>(lv, mn, Bs, g, q);
        // As[ig + 8 r][4 kq .. 4 kq + 3]
#pragma unroll
        for (int r = 0; r < kVecRows; ++r) {
            typedef T T4 __attribute__((ext_vector_type(4)));
            T4 x;
            x[0] = MF<T>::cvt(av[r].x);
            x[1] = MF<T>::cvt(av[r].y);
            x[2] = MF<T>::cvt(av[r].z);
            x[3] = MF<T>::cvt(av[r].w);
            *reinterpret_cast<T4*>(&As[ig + 8 * r][4 * kq]) = x;
        }
        const float s = msc[mycol];
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] *= s;
        __syncthreads();
        if (k0 + TK < kf) {
            if (PRE && k0 == 0) {
#pragma unroll
                for (int r = 0; r < kVecRows; ++r) {
                    lv[r] = lv2[r];
                    av[r] = av2[r];
                }
            } else {
                load(k0 + TK, lv, av);
            }
        }
        acc = stage_mma<T, LD>(As, Bs, wm, wn, lane, acc);
        __syncthreads();
    }
    if (b < B) {
        const float m = mrun[mycol];
        float v[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float e = epre[r];
            v[r] = (m == kNegInf || e == kNegInf) ? 0.0f : acc[r] * __expf(e + m);
        }
        if (FUSE == 2) {
#pragma unroll
            for (int u = 0; u < kPreChildren; ++u) {
                if (cpre[u] == 0) continue;
                const int64_t cbase = (int64_t)(cpre[u] >> 1) + m0 + wm * 32;
                if ((cpre[u] & 1) && mod_children) {
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        child_flows[(cbase + acc_row(r, lane)) * B + b] =
                            (mpre[u][r] == kNegInf) ? kNegInf : __logf(v[r]) - mpre[u][r];
                } else {
#pragma unroll
                    for (int r = 0; r < 16; ++r) child_flows[(cbase + acc_row(r, lane)) * B + b] = v[r];
                }
            }
        } else if (FUSE) {
            scatter_children(child_flows, nmars, ch_nodes + (int64_t)j * Kc, Kc, m0 + wm * 32, lane, b, B, mod_children,
                             v, FUSE == 1 ? cfirst : nullptr);
        } else {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t o = (obase + acc_row(r, lane)) * B + b;
                eflows[o] = accumulate ? eflows[o] + v[r] : v[r];
            }
        }
    }
}

// Forward with the reduction split over workgroups (16-byte RUN form, 64 x 64 tiles): for layers
// with few tiles and long child lists (an HMM's single 1024-row node block over 1024 children is
// 64 tiles of 8 dependent stages), workgroup (tile, split s) reduces stages [s P, s P + P) and
// stores its partial sum (relative to its own column max) and the max into the workspace;
// sum_fwd_split_combine then forms node = log(sum_s acc_s exp(m_s - M)) + M, M = max_s m_s.  The
// combine is a kernel of its own: a last-arriving workgroup would need a device-scope fence per
// tile, which on this chip writes back the XCD's L2 and costs more than the launch.
// Workspace: ws_acc [nsplit][n_blocks * bs][B] floats, ws_max [nsplit][n_blocks * mtiles][B] floats.
template <typename T>
__global__ __launch_bounds__(256) void sum_fwd_vec_split(float* __restrict__ node_mars, const float* __restrict__ emars,
                                                         const float* __restrict__ params, const int* __restrict__ nids,
                                                         const int* __restrict__ cids, const int* __restrict__ pids,
                                                         const int* __restrict__ nch, int K, int64_t B, int mtiles,
                                                         int btiles, int nsplit, float* __restrict__ ws_acc,
                                                         float* __restrict__ ws_max) {
    constexpr int TM = 64, TN = 64, LD = TK + MF<T>::PAD;
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    __shared__ float red[kVecGroups][64], mrun[TN], msc[TN];

    const int u = xcd_tile(blockIdx.x, gridDim.x);
    const int tile = u / nsplit, sp = u % nsplit;
    const int m0 = (tile % mtiles) * TM;
    const int j = tile / (mtiles * btiles);
    const int64_t b0 = (int64_t)((tile / mtiles) % btiles) * TN;
    const int bs = mtiles * TM;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w % 2, wn = w / 2;
    const int n = nch[j];
    const int nst = (n + TK - 1) / TK, per = (nst + nsplit - 1) / nsplit;
    const int kbeg = sp * per * TK;
    const int kend = min(n, (sp + 1) * per * TK);
    const int* crow = cids + (int64_t)j * K;
    const int* prow = pids + (int64_t)j * K;
    const int g = tid / 16, q = tid % 16;
    const int64_t bq = b0 + 4 * q;
    const bool bok = bq < B;
    const int64_t bqc = bok ? bq : B - 4;
    const int iq = tid % 16, kg = tid / 16;
    const int mycol = wn * 32 + (lane & 31);
    if (tid < TN) mrun[tid] = kNegInf;
    floatx16 acc = {0};

    float4 ev[kVecRows], av[kVecRows];
    auto load = [&](int k0) {
        const int64_t c0 = crow[k0], p0 = prow[k0];
#pragma unroll
        for (int r = 0; r < kVecRows; ++r) {
            const int k = 8 * g + r;
            const int kc = (k0 + k < kend) ? k : 0;
            ev[r] = *reinterpret_cast<const float4*>(emars + (c0 + kc) * B + bqc);
            if (!(k0 + k < kend && bok)) ev[r] = make_float4(kNegInf, kNegInf, kNegInf, kNegInf);
        }
#pragma unroll
        for (int r = 0; r < kVecRows; ++r) {
            const int k = 8 * kg + r;
            const int kc = (k0 + k < kend) ? k : 0;
            av[r] = *reinterpret_cast<const float4*>(params + p0 + (int64_t)kc * bs + m0 + 4 * iq);
            if (!(k0 + k < kend)) av[r] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    };
    if (kbeg < kend) load(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        float mn[4];
        vec_stage_max(ev, red, mrun, msc, g, q, mn);
        vec_stage_b<T, LD>(ev, mn, Bs, g, q);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float x[8];
#pragma unroll
            for (int r = 0; r < kVecRows; ++r)
                x[r] = c == 0 ? av[r].x : c == 1 ? av[r].y : c == 2 ? av[r].z : av[r].w;
            Pack8<T>::store(&As[4 * iq + c][8 * kg], x);
        }
        const float s = msc[mycol];
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] *= s;
        __syncthreads();
        if (k0 + TK < kend) load(k0 + TK);
        acc = stage_mma<T, LD>(As, Bs, wm, wn, lane, acc);
        __syncthreads();
    }
    // this split's partial
    const int64_t rows = (int64_t)gridDim.x / nsplit / btiles * TM;   // n_blocks * bs
    const int64_t b = b0 + mycol;
    const int64_t row0 = (int64_t)j * bs + m0 + wm * 32;
    if (b < B) {
#pragma unroll
        for (int r = 0; r < 16; ++r) ws_acc[((int64_t)sp * rows + row0 + acc_row(r, lane)) * B + b] = acc[r];
        if (wm == 0 && lane < 32)
            ws_max[((int64_t)sp * (rows / TM) + (int64_t)j * mtiles + m0 / TM) * B + b] = mrun[mycol];
    }
}

// Largest split count of the split-reduction kernels.
constexpr int kMaxSplits = 16;

// Log-sum-exp of the nsplit partials (running max, sum scaled to it) of one row at the four
// columns b .. b + 3: every partial is loaded before any is used, so a thread waits for one round
// trip of 16-byte loads.  M = -inf when every split is empty.
__device__ __forceinline__ void split_parts4(const float* __restrict__ ws_acc, const float* __restrict__ ws_max,
                                             int nsplit, int64_t rows, int64_t row, int64_t B, int64_t b, float M[4],
                                             float S[4]) {
    float4 mx[kMaxSplits], ac[kMaxSplits];
    const int64_t mrow = row / 64;
#pragma unroll
    for (int s = 0; s < kMaxSplits; ++s) {
        if (s < nsplit) {
            mx[s] = *reinterpret_cast<const float4*>(ws_max + ((int64_t)s * (rows / 64) + mrow) * B + b);
            ac[s] = *reinterpret_cast<const float4*>(ws_acc + ((int64_t)s * rows + row) * B + b);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float m = kNegInf;
#pragma unroll
        for (int s = 0; s < kMaxSplits; ++s)
            if (s < nsplit) m = fmaxf(m, c == 0 ? mx[s].x : c == 1 ? mx[s].y : c == 2 ? mx[s].z : mx[s].w);
        float sum = 0.0f;
        if (m != kNegInf) {
#pragma unroll
            for (int s = 0; s < kMaxSplits; ++s) {
                if (s < nsplit) {
                    const float ms = c == 0 ? mx[s].x : c == 1 ? mx[s].y : c == 2 ? mx[s].z : mx[s].w;
                    const float a = c == 0 ? ac[s].x : c == 1 ? ac[s].y : c == 2 ? ac[s].z : ac[s].w;
                    sum += (ms == kNegInf ? 0.0f : __expf(ms - m)) * a;
                }
            }
        }
        M[c] = m;
        S[c] = sum;
    }
}

// node_mars[nids[j] + i][b] from the partials of sum_fwd_vec_split; one thread per (row, four
// columns), B % 4 == 0.
__global__ __launch_bounds__(256) void sum_fwd_split_combine(float* __restrict__ node_mars, const int* __restrict__ nids,
                                                             const float* __restrict__ ws_acc,
                                                             const float* __restrict__ ws_max, int nsplit, int bs,
                                                             int64_t rows, int64_t B) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nq = B / 4;
    if (t >= rows * nq) return;
    const int64_t row = t / nq, b = 4 * (t % nq);
    float M[4], S[4];
    split_parts4(ws_acc, ws_max, nsplit, rows, row, B, b, M, S);
    const int64_t j = row / bs, i = row % bs;
    float4 o;
    o.x = (S[0] > 0.0f && M[0] != kNegInf) ? __logf(S[0]) + M[0] : kNegInf;
    o.y = (S[1] > 0.0f && M[1] != kNegInf) ? __logf(S[1]) + M[1] : kNegInf;
    o.z = (S[2] > 0.0f && M[2] != kNegInf) ? __logf(S[2]) + M[2] : kNegInf;
    o.w = (S[3] > 0.0f && M[3] != kNegInf) ? __logf(S[3]) + M[3] : kNegInf;
    *reinterpret_cast<float4*>(node_mars + ((int64_t)nids[j] + i) * B + b) = o;
}

// Element flows with the reduction over parent rows split over workgroups (16-byte form, parent
// blocks of >= TK rows, child blocks of >= 64 rows), the counterpart of sum_fwd_vec_split for the
// flow backward of few-tile layers; sum_ele_split_combine then forms each child row's flow and
// stores it into the children of the product row (the fused product backward).  Workspace as
// sum_fwd_vec_split, over child rows.
template <typename T>
__global__ __launch_bounds__(256) void sum_ele_vec_split(const float* nflows, const float* __restrict__ nmars,
                                                         const float* __restrict__ params, const int* __restrict__ parids,
                                                         const int* __restrict__ parpids, const int* __restrict__ npar,
                                                         int KP, int lbs, int64_t B, int mtiles, int btiles,
                                                         bool modified, int nsplit, float* __restrict__ ws_acc,
                                                         float* __restrict__ ws_max) {
    constexpr int TM = 64, TN = 64, LD = TK + MF<T>::PAD;
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    __shared__ float red[kVecGroups][64], mrun[TN], msc[TN];

    const int u = xcd_tile(blockIdx.x, gridDim.x);
    const int tile = u / nsplit, sp = u % nsplit;
    const int m0 = (tile % mtiles) * TM;
    const int j = tile / (mtiles * btiles);
    const int64_t b0 = (int64_t)((tile / mtiles) % btiles) * TN;
    const int bs = 1 << lbs;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w % 2, wn = w / 2;
    const int kf = npar[j] << lbs;
    const int nst = kf / TK, per = (nst + nsplit - 1) / nsplit;
    const int kbeg = sp * per * TK;
    const int kend = min(kf, (sp + 1) * per * TK);
    const int* parrow = parids + (int64_t)j * KP;
    const int* pprow = parpids + (int64_t)j * KP;
    const int g = tid / 16, q = tid % 16;
    const int64_t bq = b0 + 4 * q;
    const bool bok = bq < B;
    const int64_t bqc = bok ? bq : B - 4;
    const int kq = tid % 32, ig = tid / 32;
    const int mycol = wn * 32 + (lane & 31);
    if (tid < TN) mrun[tid] = kNegInf;
    floatx16 acc = {0};

    float4 lv[kVecRows], av[kVecRows];
    auto load = [&](int k0) {
        const int kb = k0 >> lbs, k_in = k0 & (bs - 1);
        const int64_t rbase = (int64_t)parrow[kb] + k_in + 8 * g;
#pragma unroll
        for (int r = 0; r < kVecRows; ++r) {
            const int64_t o = (rbase + r) * B + bqc;
            float4 f = *reinterpret_cast<const float4*>(nflows + o);
            if (!modified) {
                const float4 m = *reinterpret_cast<const float4*>(nmars + o);
                f.x = m.x == kNegInf ? kNegInf : __logf(f.x) - m.x;
                f.y = m.y == kNegInf ? kNegInf : __logf(f.y) - m.y;
                f.z = m.z == kNegInf ? kNegInf : __logf(f.z) - m.z;
                f.w = m.w == kNegInf ? kNegInf : __logf(f.w) - m.w;
            }
            lv[r] = bok ? f : make_float4(kNegInf, kNegInf, kNegInf, kNegInf);
        }
        const float* pp = params + (int64_t)pprow[kb] + (int64_t)(m0 + ig) * bs + k_in + 4 * kq;
#pragma unroll
        for (int r = 0; r < kVecRows; ++r) av[r] = *reinterpret_cast<const float4*>(pp + (int64_t)(8 * r) * bs);
    };
    if (kbeg < kend) load(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        float mn[4];
        vec_stage_max(lv, red, mrun, msc, g, q, mn);
        vec_stage_b<T, LD>(lv, mn, Bs, g, q);
#pragma unroll
        for (int r = 0; r < kVecRows; ++r) {
            typedef T T4 __attribute__((ext_vector_type(4)));
            T4 x;
            x[0] = MF<T>::cvt(av[r].x);
            x[1] = MF<T>::cvt(av[r].y);
            x[2] = MF<T>::cvt(av[r].z);
            x[3] = MF<T>::cvt(av[r].w);
            *reinterpret_cast<T4*>(&As[ig + 8 * r][4 * kq]) = x;
        }
        const float s = msc[mycol];
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] *= s;
        __syncthreads();
        if (k0 + TK < kend) load(k0 + TK);
        acc = stage_mma<T, LD>(As, Bs, wm, wn, lane, acc);
        __syncthreads();
    }
    const int64_t rows = (int64_t)gridDim.x / nsplit / btiles * TM;   // n_cblocks * cbs
    const int64_t b = b0 + mycol;
    const int64_t row0 = (int64_t)j * mtiles * TM + m0 + wm * 32;
    if (b < B) {
#pragma unroll
        for (int r = 0; r < 16; ++r) ws_acc[((int64_t)sp * rows + row0 + acc_row(r, lane)) * B + b] = acc[r];
        if (wm == 0 && lane < 32)
            ws_max[((int64_t)sp * (rows / TM) + (int64_t)j * mtiles + m0 / TM) * B + b] = mrun[mycol];
    }
}

// Child-row flows from the partials of sum_ele_vec_split, written into the product's children
// (ch_nodes as pjamd_sum_backward_ele_to_children); one thread per (child row, four columns),
// B % 4 == 0.
__global__ __launch_bounds__(256) void sum_ele_split_combine(float* __restrict__ child_flows,
                                                             const float* __restrict__ nmars,
                                                             const float* __restrict__ emars,
                                                             const int* __restrict__ chids,
                                                             const int* __restrict__ ch_nodes, int Kc,
                                                             bool mod_children, const float* __restrict__ ws_acc,
                                                             const float* __restrict__ ws_max, int nsplit, int cbs,
                                                             int64_t rows, int64_t B) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nq = B / 4;
    if (t >= rows * nq) return;
    const int64_t row = t / nq, b = 4 * (t % nq);
    const int64_t j = row / cbs, i = row % cbs;
    const float4 e = *reinterpret_cast<const float4*>(emars + ((int64_t)chids[j] + i) * B + b);
    float M[4], S[4];
    split_parts4(ws_acc, ws_max, nsplit, rows, row, B, b, M, S);
    const float ev[4] = {e.x, e.y, e.z, e.w};
    float v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = (M[c] == kNegInf || ev[c] == kNegInf) ? 0.0f : S[c] * __expf(ev[c] + M[c]);
    const int* crow = ch_nodes + j * Kc;
    for (int k = 0; k < Kc; ++k) {
        const int code = crow[k];
        if (code == 0) continue;
        const int64_t o = ((int64_t)(code >> 1) + i) * B + b;
        float4 w = make_float4(v[0], v[1], v[2], v[3]);
        if ((code & 1) && mod_children) {
            const float4 m = *reinterpret_cast<const float4*>(nmars + o);
            w.x = (m.x == kNegInf) ? kNegInf : __logf(v[0]) - m.x;
            w.y = (m.y == kNegInf) ? kNegInf : __logf(v[1]) - m.y;
            w.z = (m.z == kNegInf) ? kNegInf : __logf(v[2]) - m.z;
            w.w = (m.w == kNegInf) ? kNegInf : __logf(v[3]) - m.w;
        }
        *reinterpret_cast<float4*>(child_flows + o) = w;
    }
}

// Forward with 16-byte loads for node blocks of 32 rows (32 x 128 tiles) and for child runs
// shorter than a stage (runs of >= 16 consecutive element rows with parameters bs apart, e.g. the
// child blocks of 32 of a block-size-32 layer).  The block's index row is staged in LDS (in chunks
// of kFwdIdx entries); each thread reads the run start of its own rows there, so every load is
// unconditional (row clamped to a real child, value masked afterwards) and a stage is one round
// trip of 16-byte loads.  Requirements (checked by the launcher): 16-bit operands, B % 4 == 0,
// 16-byte-aligned parameter rows, child runs >= 16.
//
// Thread layout, TN = 4096 / TM columns: element rows: column quad q = tid % (TN / 4), row group
// g of RPG = TK * TN / 1024 consecutive children; parameters: node-row quad iq = tid % (TM / 4),
// child group kg of KPG = TM / 8 consecutive children.
constexpr int kFwdIdx = 512;

template <typename T, int TM, bool QUEUE>
__global__ __launch_bounds__(256) void sum_fwd_vec_blocks(float* __restrict__ node_mars, const float* __restrict__ emars,
                                                          const float* __restrict__ params, const int* __restrict__ nids,
                                                          const int* __restrict__ cids, const int* __restrict__ pids,
                                                          const int* __restrict__ nch, int K, int64_t B, int mtiles,
                                                          int btiles, int* __restrict__ queue, int n_items) {
    constexpr int TN = 4096 / TM, WM = TM / 32;
    constexpr int QN = TN / 4, GR = 256 / QN, RPG = TK / GR;
    constexpr int QM = TM / 4, KPG = TM / 8;
    constexpr int LD = TK + MF<T>::PAD;
    static_assert((RPG == 8 || RPG == 16) && (KPG == 4 || KPG == 8), "row groups");
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    __shared__ float red[GR][TN], mrun[TN], msc[TN];
    __shared__ int sc[kFwdIdx], sp[kFwdIdx];
    __shared__ int s_item;

    for (int it = 0;; ++it) {
    int t;
    if constexpr (QUEUE) {
        // work queue: the next tile in dispatch order (heaviest node blocks first)
        __syncthreads();                 // the previous tile is done with the shared arrays
        if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
        __syncthreads();
        t = s_item;
        if (t >= n_items) break;
    } else {
        if (it > 0) break;
        t = xcd_tile(blockIdx.x, gridDim.x);
    }
    const int m0 = (t % mtiles) * TM;
    const int j = t / (mtiles * btiles);
    const int64_t b0 = (int64_t)((t / mtiles) % btiles) * TN;
    const int bs = mtiles * TM;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w % WM, wn = w / WM;
    const int n = nch[j];
    const int q = tid % QN, g = tid / QN;
    const int64_t bq = b0 + 4 * q;
    const bool bok = bq < B;                          // B % 4 == 0: a quad is all in or all out
    const int64_t bqc = bok ? bq : B - 4;
    const int iq = tid % QM, kg = tid / QM;
    const int mycol = wn * 32 + (lane & 31);
    if (tid < TN) mrun[tid] = kNegInf;
    floatx16 acc = {0};

    float4 ev[RPG], av[KPG];
    bool eok[RPG], aok[KPG];
    // the children of this chunk are [cb, cb + cn); k below counts from cb
    int cn = 0;
    auto load = [&](int k0) {
        // element rows k0 + RPG g + r: one run (RPG divides the run length), clamped to the last
        // real child so that every address is a real row
        const int ke = k0 + RPG * g;
        const int kec = ke < cn ? ke : 0;
        const int64_t c0 = sc[kec];
        const int last = cn - 1 - kec;                // rows of the run past this are not real
#pragma unroll
        for (int r = 0; r < RPG; ++r) {
            const int rr = r < last ? r : (last > 0 ? last : 0);
            ev[r] = *reinterpret_cast<const float4*>(emars + (c0 + rr) * B + bqc);
            eok[r] = (ke + r < cn) && bok;
        }
        const int ka = k0 + KPG * kg;
        const int kac = ka < cn ? ka : 0;
        const int64_t p0 = sp[kac];
        const int lasta = cn - 1 - kac;
#pragma unroll
        for (int r = 0; r < KPG; ++r) {
            const int rr = r < lasta ? r : (lasta > 0 ? lasta : 0);
            av[r] = *reinterpret_cast<const float4*>(params + p0 + (int64_t)rr * bs + m0 + 4 * iq);
            aok[r] = ka + r < cn;
        }
    };
    // index rows longer than kFwdIdx are staged in chunks of kFwdIdx (a multiple of TK, so runs
    // never straddle chunks); the running max and the accumulator carry over
    for (int cb = 0; cb < n; cb += kFwdIdx) {
        cn = (n - cb < kFwdIdx) ? n - cb : kFwdIdx;
        const int ck = (K - cb < kFwdIdx) ? K - cb : kFwdIdx;
        __syncthreads();
        for (int u = tid; u < ck; u += 256) {
            sc[u] = cids[(int64_t)j * K + cb + u];
            sp[u] = pids[(int64_t)j * K + cb + u];
        }
        __syncthreads();
        load(0);
        for (int k0 = 0; k0 < cn; k0 += TK) {
            float mx[4] = {kNegInf, kNegInf, kNegInf, kNegInf};
#pragma unroll
            for (int r = 0; r < RPG; ++r) {
                if (!eok[r]) ev[r] = make_float4(kNegInf, kNegInf, kNegInf, kNegInf);
                mx[0] = fmaxf(mx[0], ev[r].x);
                mx[1] = fmaxf(mx[1], ev[r].y);
                mx[2] = fmaxf(mx[2], ev[r].z);
                mx[3] = fmaxf(mx[3], ev[r].w);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) red[g][4 * q + c] = mx[c];
            __syncthreads();
            if (tid < TN) {
                float x = red[0][tid];
#pragma unroll
                for (int i = 1; i < GR; ++i) x = fmaxf(x, red[i][tid]);
                const float mo = mrun[tid];
                const float mnew = fmaxf(mo, x);
                msc[tid] = (mnew == kNegInf) ? 1.0f : __expf(mo - mnew);
                mrun[tid] = mnew;
            }
            __syncthreads();
            // B tile: Bs[4 q + c][RPG g + r] = exp(ev[r][c] - max[c])
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float ns = neg_shift2(mrun[4 * q + c]);   // exp(x - max); 0 for -inf
#pragma unroll
                for (int h = 0; h < RPG / 8; ++h) {
                    float e[8];
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
                        const float4 v = ev[8 * h + r];
                        const float x = c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
                        e[r] = op_exp2(fmaf(x, kLog2e, ns));
                    }
                    Pack8<T>::store(&Bs[4 * q + c][RPG * g + 8 * h], e);
                }
            }
            // A tile: As[4 iq + c][KPG kg + r] = params of node row m0 + 4 iq + c, child KPG kg + r
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                typedef T TV __attribute__((ext_vector_type(KPG)));
                TV x;
#pragma unroll
                for (int r = 0; r < KPG; ++r) {
                    const float4 v = av[r];
                    const float y = c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
                    x[r] = MF<T>::cvt(aok[r] ? y : 0.0f);
                }
                *reinterpret_cast<TV*>(&As[4 * iq + c][KPG * kg]) = x;
            }
            const float s = msc[mycol];
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] *= s;
            __syncthreads();
            if (k0 + TK < cn) load(k0 + TK);
            acc = stage_mma<T, LD>(As, Bs, wm, wn, lane, acc);
            __syncthreads();
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
    if constexpr (QUEUE) {
        // the last workgroup to leave resets the queue for the next launch
        if (threadIdx.x == 0) {
            __threadfence();
            if (atomicAdd(queue + 1, 1) == (int)gridDim.x - 1) {
                atomicExch(queue, 0);
                atomicExch(queue + 1, 0);
            }
        }
    }
}

// Element flows with 16-byte loads for the shapes sum_ele_vec does not take: child blocks of 32
// rows (32 x 128 tiles) and parent blocks smaller than a stage (a stage of TK flattened parent
// rows then spans several parent blocks).  The parent-block indices of the child block (at most
// kEleIdx) are staged in LDS once; every thread finds the parent block of its own rows there, so
// each load is unconditional (index clamped to a real entry, value masked afterwards) and a stage
// is one round trip of 16-byte loads.  Node flows must already hold log(flow) - mar.
// Requirements (checked by the launcher): 16-bit operands, B % 4 == 0, parent blocks of >= 16 rows,
// 16-byte-aligned parameter rows, n_pars entries <= kEleIdx.  FUSE as sum_ele_mfma (0 or 1).
//
// Thread layout, TN = 4096 / TM columns: flow rows: column quad q = tid % (TN / 4), row group
// g = tid / (TN / 4) of RPG = TK * TN / 1024 consecutive rows (one parent block: RPG <= 16 <= bs);
// parameters: k quad kq = tid % 32 (4 consecutive flattened parent rows, one parent block since
// bs % 4 == 0), child rows ig + 8 r.
constexpr int kEleIdx = 256;

template <typename T, int TM, int FUSE>
__global__ __launch_bounds__(256) void sum_ele_vec_blocks(float* __restrict__ eflows, const float* nflows,
                                                          const float* __restrict__ nmars, const float* __restrict__ emars,
                                                          const float* __restrict__ params, const int* __restrict__ chids,
                                                          const int* __restrict__ parids, const int* __restrict__ parpids,
                                                          const int* __restrict__ npar, int KP, int lbs, int64_t B,
                                                          int mtiles, int btiles, bool accumulate, float* child_flows,
                                                          const int* __restrict__ ch_nodes, int Kc, bool mod_children) {
    constexpr int TN = 4096 / TM, WM = TM / 32;
    constexpr int QN = TN / 4, GR = 256 / QN, RPG = TK / GR, AR = TM / 8;
    constexpr int LD = TK + MF<T>::PAD;
    static_assert(RPG == 8 || RPG == 16, "row groups of 8 or 16 flow rows");
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    __shared__ float red[GR][TN], mrun[TN], msc[TN];
    __shared__ int spar[kEleIdx], spp[kEleIdx];

    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int m0 = (t % mtiles) * TM;
    const int j = t / (mtiles * btiles);
    const int64_t b0 = (int64_t)((t / mtiles) % btiles) * TN;
    const int bs = 1 << lbs;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w % WM, wn = w / WM;
    const int np = npar[j];
    const int q = tid % QN, g = tid / QN;
    const int64_t bq = b0 + 4 * q;
    const bool bok = bq < B;                          // B % 4 == 0: a quad is all in or all out
    const int64_t bqc = bok ? bq : B - 4;
    const int kq = tid % 32, ig = tid / 32;
    const int mycol = wn * 32 + (lane & 31);
    for (int u = tid; u < KP; u += 256) {
        spar[u] = parids[(int64_t)j * KP + u];
        spp[u] = parpids[(int64_t)j * KP + u];
    }
    if (tid < TN) mrun[tid] = kNegInf;
    floatx16 acc = {0};
    const int kf = np << lbs;

    // element marginals of the epilogue, independent of the reduction
    const int64_t b = b0 + mycol;
    const int64_t obase = (int64_t)chids[j] + m0 + wm * 32;
    float epre[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) epre[r] = (b < B) ? emars[(obase + acc_row(r, lane)) * B + b] : kNegInf;
    __syncthreads();

    float4 lv[RPG], av[AR];
    auto load = [&](int k0) {
        const int kr = k0 + RPG * g;                  // first flattened row of this thread's group
        const bool rok = kr < kf;
        const int krc = rok ? kr : 0;
        const int64_t rbase = (int64_t)spar[krc >> lbs] + (krc & (bs - 1));
#pragma unroll
        for (int r = 0; r < RPG; ++r) lv[r] = *reinterpret_cast<const float4*>(nflows + (rbase + r) * B + bqc);
#pragma unroll
        for (int r = 0; r < RPG; ++r)
            if (!(rok && bok)) lv[r] = make_float4(kNegInf, kNegInf, kNegInf, kNegInf);
        const int ka = k0 + 4 * kq;
        const bool aok = ka < kf;
        const int kac = aok ? ka : 0;
        const float* pp = params + (int64_t)spp[kac >> lbs] + (int64_t)(m0 + ig) * bs + (kac & (bs - 1));
#pragma unroll
        for (int r = 0; r < AR; ++r) av[r] = *reinterpret_cast<const float4*>(pp + (int64_t)(8 * r) * bs);
#pragma unroll
        for (int r = 0; r < AR; ++r)
            if (!aok) av[r] = make_float4(0.f, 0.f, 0.f, 0.f);
    };
    if (kf > 0) load(0);
    for (int k0 = 0; k0 < kf; k0 += TK) {
        // column max of the stage, folded into the running max
        float mx[4] = {kNegInf, kNegInf, kNegInf, kNegInf};
#pragma unroll
        for (int r = 0; r < RPG; ++r) {
            mx[0] = fmaxf(mx[0], lv[r].x);
            mx[1] = fmaxf(mx[1], lv[r].y);
            mx[2] = fmaxf(mx[2], lv[r].z);
            mx[3] = fmaxf(mx[3], lv[r].w);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) red[g][4 * q + c] = mx[c];
        __syncthreads();
        if (tid < TN) {
            float x = red[0][tid];
#pragma unroll
            for (int i = 1; i < GR; ++i) x = fmaxf(x, red[i][tid]);
            const float mo = mrun[tid];
            const float mnew = fmaxf(mo, x);
            msc[tid] = (mnew == kNegInf) ? 1.0f : __expf(mo - mnew);
            mrun[tid] = mnew;
        }
        __syncthreads();
        // B tile: Bs[4 q + c][RPG g + r] = exp(lv[r][c] - max[c])
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float ns = neg_shift2(mrun[4 * q + c]);   // exp(x - max); 0 for -inf
#pragma unroll
            for (int h = 0; h < RPG / 8; ++h) {
                float e[8];
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const float4 v = lv[8 * h + r];
                    const float x = c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
                    e[r] = op_exp2(fmaf(x, kLog2e, ns));
                }
                Pack8<T>::store(&Bs[4 * q + c][RPG * g + 8 * h], e);
            }
        }
        // A tile: As[ig + 8 r][4 kq .. 4 kq + 3]
#pragma unroll
        for (int r = 0; r < AR; ++r) {
            typedef T T4 __attribute__((ext_vector_type(4)));
            T4 x;
            x[0] = MF<T>::cvt(av[r].x);
            x[1] = MF<T>::cvt(av[r].y);
            x[2] = MF<T>::cvt(av[r].z);
            x[3] = MF<T>::cvt(av[r].w);
            *reinterpret_cast<T4*>(&As[ig + 8 * r][4 * kq]) = x;
        }
        const float s = msc[mycol];
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] *= s;
        __syncthreads();
        if (k0 + TK < kf) load(k0 + TK);
        acc = stage_mma<T, LD>(As, Bs, wm, wn, lane, acc);
        __syncthreads();
    }
    if (b < B) {
        const float m = mrun[mycol];
        float v[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float e = epre[r];
            v[r] = (m == kNegInf || e == kNegInf) ? 0.0f : acc[r] * __expf(e + m);
        }
        if (FUSE) {
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
                eflows[o] = accumulate ? eflows[o] + v[r] : v[r];
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Parameter flows of one tile = (node block j, TM parent rows from m0) x (TN children from kt0),
// reduced over the batch range [bstart, bend).  Both operand tiles are loaded with lanes along
// the batch.  Shared by the per-layer and the batched kernels below.
template <typename T, int TM, int TN, typename PF, bool VEC>
__device__ __forceinline__ void par_tile(PF* __restrict__ pflows, const float* __restrict__ nflows,
                                         const float* __restrict__ nmars, const float* __restrict__ emars,
                                         const float* __restrict__ params, const int* __restrict__ nids,
                                         const int* __restrict__ cids, const int* __restrict__ pids,
                                         const int* __restrict__ pfids, const int* __restrict__ nch, int K, int64_t B,
                                         int j, int m0, int kt0, int64_t bstart, int64_t bend, bool modified,
                                         int atomic_out, float sign, const int* __restrict__ copies = nullptr,
                                         int n_copies = 1) {
    constexpr int WM = TM / 32, WN = TN / 32;
    static_assert(WM * WN == 4, "four wavefronts per workgroup");
    constexpr int LD = TK + MF<T>::PAD;
    constexpr int RG = 256 / TK;        // rows per pass of a tile load (lanes run along the batch)
    constexpr int PX = TM / RG, PY = TN / RG;
    __shared__ __attribute__((aligned(16))) T As[TM][LD];
    __shared__ __attribute__((aligned(16))) T Bs[TN][LD];
    constexpr int RGV = 256 / (TK / 4);  // VEC: rows per pass (lanes run along the batch, 4 values each)
    __shared__ float xred[VEC ? RGV : RG][TK];
    __shared__ float sh[TK];
    __shared__ int64_t sy[TN];          // element row offsets of the tile's children (-1: padding)

    const int n = nch[j];
    if (kt0 >= n) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w % WM, wn = w / WM;
    floatx16 acc = {0};
    // Tied copies (rows copies[r * PJAMD_PAR_DESC_FIELDS]: block offset, index-row offset and
    // element row relative to copy 0, same parameter and flow slots) add into one accumulator.
    for (int r = 0; r < n_copies; ++r) {
    int64_t nrow0 = (int64_t)nids[j] + m0;
    const int* crow = cids + (int64_t)j * K;
    const float* em = emars;
    if (r > 0) {
        const int* R = copies + (int64_t)r * PJAMD_PAR_DESC_FIELDS;
        nrow0 = (int64_t)nids[(int64_t)R[4] - copies[4] + j] + m0;
        crow = cids + ((int64_t)R[5] - copies[5]) + (int64_t)j * K;
        em = emars + ((int64_t)R[6] - copies[6]) * B;
    }

    if (tid < TN) sy[tid] = (kt0 + tid < n) ? (int64_t)crow[kt0 + tid] * B : -1;
    __syncthreads();

    if constexpr (VEC) {
        // 16-byte loads (B % 4 == 0, batch range bounds multiples of 4): thread = (row group
        // rg = tid / 32, batch quad xq = tid % 32); 16 float4 per stage instead of 64 dwords.
        constexpr int PXV = TM / RGV, PYV = TN / RGV;
        const int xq = tid % 32, rg = tid / 32;
        float4 xv[PXV], yv[PYV];
        const float4 ninf = make_float4(kNegInf, kNegInf, kNegInf, kNegInf);
        auto load = [&](int64_t b0) {
            const int64_t b = b0 + 4 * xq;
            const bool ok = b < bend;
            const int64_t bc = ok ? b : bstart;
#pragma unroll
            for (int q = 0; q < PXV; ++q) {
                const int64_t o = (nrow0 + rg + (int64_t)q * RGV) * B + bc;
                float4 f = *reinterpret_cast<const float4*>(nflows + o);
                if (!modified) {
                    const float4 m = *reinterpret_cast<const float4*>(nmars + o);
                    f.x = m.x == kNegInf ? kNegInf : __logf(f.x) - m.x;
                    f.y = m.y == kNegInf ? kNegInf : __logf(f.y) - m.y;
                    f.z = m.z == kNegInf ? kNegInf : __logf(f.z) - m.z;
                    f.w = m.w == kNegInf ? kNegInf : __logf(f.w) - m.w;
                }
                xv[q] = ok ? f : ninf;
            }
#pragma unroll
            for (int q = 0; q < PYV; ++q) {
                const int64_t yb = sy[rg + q * RGV];
                const float4 y = *reinterpret_cast<const float4*>(em + (yb >= 0 ? yb : 0) + bc);
                yv[q] = (ok && yb >= 0) ? y : ninf;
            }
        };
        load(bstart);
        for (int64_t b0 = bstart; b0 < bend; b0 += TK) {
            float m4[4] = {kNegInf, kNegInf, kNegInf, kNegInf};
#pragma unroll
            for (int q = 0; q < PXV; ++q) {
                m4[0] = fmaxf(m4[0], xv[q].x);
                m4[1] = fmaxf(m4[1], xv[q].y);
                m4[2] = fmaxf(m4[2], xv[q].z);
                m4[3] = fmaxf(m4[3], xv[q].w);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) xred[rg][4 * xq + c] = m4[c];
            __syncthreads();
            if (tid < TK) {
                float v = xred[0][tid];
#pragma unroll
                for (int g = 1; g < RGV; ++g) v = fmaxf(v, xred[g][tid]);
                sh[tid] = neg_shift2(v);
            }
            __syncthreads();
            float s4[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) s4[c] = sh[4 * xq + c];
            typedef T T4 __attribute__((ext_vector_type(4)));
            // -inf entries (padding, zero flows, impossible children) give exp2(-inf) = 0.
#pragma unroll
            for (int q = 0; q < PXV; ++q) {
                const float x[4] = {xv[q].x, xv[q].y, xv[q].z, xv[q].w};
                T4 o;
#pragma unroll
                for (int c = 0; c < 4; ++c) o[c] = MF<T>::cvt(op_exp2(fmaf(x[c], kLog2e, s4[c])));
                *reinterpret_cast<T4*>(&As[rg + q * RGV][4 * xq]) = o;
            }
#pragma unroll
            for (int q = 0; q < PYV; ++q) {
                const float y[4] = {yv[q].x, yv[q].y, yv[q].z, yv[q].w};
                T4 o;
#pragma unroll
                for (int c = 0; c < 4; ++c) o[c] = MF<T>::cvt(op_exp2(fmaf(y[c], kLog2e, -s4[c])));
                *reinterpret_cast<T4*>(&Bs[rg + q * RGV][4 * xq]) = o;
            }
            __syncthreads();
            if (b0 + TK < bend) load(b0 + TK);
            acc = stage_mma<T, LD>(As, Bs, wm, wn, lane, acc);
        }
    } else {
    const int xb = tid % TK, r0 = tid / TK;
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
            yv[q] = (ok && yb >= 0) ? em[yb + b] : kNegInf;
        }
    };

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
        const float s = neg_shift2(sh[xb]);   // exp(x - max), exp(y + max); 0 for -inf entries
#pragma unroll
        for (int q = 0; q < PX; ++q) As[r0 + q * RG][xb] = MF<T>::cvt(op_exp2(fmaf(xv[q], kLog2e, s)));
#pragma unroll
        for (int q = 0; q < PY; ++q) Bs[r0 + q * RG][xb] = MF<T>::cvt(op_exp2(fmaf(yv[q], kLog2e, -s)));
        __syncthreads();
        if (b0 + TK < bend) load(b0 + TK);
        acc = stage_mma<T, LD>(As, Bs, wm, wn, lane, acc);
    }
    }
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

// ---------------------------------------------------------------------------------------------
// 128 x 128 parameter-flow tiles (BF16 / FP16 operands, B % 4 == 0, block sizes >= 128).  A tile
// moves (TM + TN) B floats from L2 for TM TN B products, so 128 x 128 tiles halve both the
// L2 -> CU traffic of 64 x 64 tiles (the limit of that form: a CU keeps a bounded number of L2
// requests in flight) and the exp work per parameter.  (256 x 128 tiles, one workgroup of eight
// wavefronts per CU, measured slower: nothing overlaps a workgroup's barriers.)  Four
// wavefronts, each a 64 x 64 quadrant of four 32 x 32 accumulators, stages of TKB = 64 batch
// columns; thread (rg = tid / 16, xq = tid % 16) loads rows rg + 16 q (q < 8) of both operand
// tiles, columns 4 xq .. 4 xq + 3.  The column max is reduced inside each wavefront by lane
// shuffles and across the four through LDS.  The epilogue stages each quadrant through LDS so
// that the slots of one child (consecutive rows) are read and written as contiguous runs.
struct BigTile {
    static constexpr int TM = 128, TN = 128, NW = 4, NT = 64 * NW, RG = NT / 16;
    static constexpr int NQX = TM / RG, NQY = TN / RG;
    static constexpr int LD = TKB + 8;    // operand row: 72 BF16 / FP16 values
    static constexpr int SLD = 64 + 4;    // epilogue staging: 64 rows (+ pad) per child, floats
    static constexpr int OPS_BYTES = (TM + TN) * LD * 2;
    static_assert(NW * 32 * SLD * 4 <= OPS_BYTES, "epilogue staging fits the operand tiles");
};

template <typename T, typename PF>
__device__ __forceinline__ void par_tile_big(PF* __restrict__ pflows, const float* __restrict__ nflows,
                                             const float* __restrict__ nmars, const float* __restrict__ emars,
                                             const float* __restrict__ params, const int* __restrict__ nids,
                                             const int* __restrict__ cids, const int* __restrict__ pids,
                                             const int* __restrict__ pfids, const int* __restrict__ nch, int K,
                                             int64_t B, int j, int m0, int kt0, bool modified, int mode, float sign,
                                             const int* __restrict__ copies, int n_copies) {
    typedef BigTile G;
    constexpr int TM = G::TM, TN = G::TN, LD = G::LD, NW = G::NW, RG = G::RG, NQX = G::NQX, NQY = G::NQY;
    constexpr int SLD = G::SLD;
    static_assert(sizeof(T) == 2 && MF<T>::PAD == 8, "16-bit operands");
    typedef T T4 __attribute__((ext_vector_type(4)));
    __shared__ __attribute__((aligned(16))) char ops_raw[G::OPS_BYTES];
    __shared__ __attribute__((aligned(16))) float red[NW][TKB];
    __shared__ int64_t sy[TN];
    __shared__ int spid[TN], spf[TN];
    T* ops = reinterpret_cast<T*>(ops_raw);
    T (*As)[LD] = reinterpret_cast<T (*)[LD]>(ops);
    T (*Bs)[LD] = reinterpret_cast<T (*)[LD]>(ops + TM * LD);

    const int n = nch[j];
    if (kt0 >= n) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w & 1, wn = w >> 1;
    const int rg = tid >> 4, xq = tid & 15;
    if (tid < TN) {
        const int k = kt0 + tid < n ? kt0 + tid : kt0;
        spid[tid] = pids[(int64_t)j * K + k];
        spf[tid] = pfids[(int64_t)j * K + k];
    }
    floatx16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = floatx16{0};
    const float4 ninf = make_float4(kNegInf, kNegInf, kNegInf, kNegInf);

    // Tied copies (rows copies[r * PJAMD_PAR_DESC_FIELDS]) add into the same accumulators.
    for (int r = 0; r < n_copies; ++r) {
        int64_t nrow0 = (int64_t)nids[j] + m0;
        const int* crow = cids + (int64_t)j * K;
        const float* em = emars;
        if (r > 0) {
            const int* R = copies + (int64_t)r * PJAMD_PAR_DESC_FIELDS;
            nrow0 = (int64_t)nids[(int64_t)R[4] - copies[4] + j] + m0;
            crow = cids + ((int64_t)R[5] - copies[5]) + (int64_t)j * K;
            em = emars + ((int64_t)R[6] - copies[6]) * B;
        }
        __syncthreads();   // the previous copy's child rows are no longer read
        // Padding children (k >= n) read child kt0's row: their columns are never written.
        if (tid < TN) sy[tid] = (int64_t)crow[kt0 + tid < n ? kt0 + tid : kt0] * B;
        __syncthreads();
        const float* xrow = nflows + nrow0 * B;
        const float* mrow = nmars + nrow0 * B;
        float4 xv[NQX], yv[NQY];
        auto load = [&](int64_t b0) {
            const int64_t b = b0 + 4 * xq;
            const int64_t bc = b < B ? b : B - 4;   // clamped: the stage's tail columns are masked
#pragma unroll
            for (int q = 0; q < NQX; ++q) {
                const int64_t o = (int64_t)(rg + RG * q) * B + bc;
                float4 f = *reinterpret_cast<const float4*>(xrow + o);
                if (!modified) {
                    const float4 m = *reinterpret_cast<const float4*>(mrow + o);
                    f.x = m.x == kNegInf ? kNegInf : __logf(f.x) - m.x;
                    f.y = m.y == kNegInf ? kNegInf : __logf(f.y) - m.y;
                    f.z = m.z == kNegInf ? kNegInf : __logf(f.z) - m.z;
                    f.w = m.w == kNegInf ? kNegInf : __logf(f.w) - m.w;
                }
                xv[q] = f;
            }
#pragma unroll
            for (int q = 0; q < NQY; ++q) yv[q] = *reinterpret_cast<const float4*>(em + sy[rg + RG * q] + bc);
        };
        load(0);
        for (int64_t b0 = 0; b0 < B; b0 += TKB) {
            // Columns past B: -inf node flows give a column max of -inf, hence zero operands on
            // both sides (the child side through the floored shift).
            if (b0 + 4 * xq >= B) {
#pragma unroll
                for (int q = 0; q < NQX; ++q) xv[q] = ninf;
            }
            float m4[4] = {kNegInf, kNegInf, kNegInf, kNegInf};
#pragma unroll
            for (int q = 0; q < NQX; ++q) {
                m4[0] = fmaxf(m4[0], xv[q].x);
                m4[1] = fmaxf(m4[1], xv[q].y);
                m4[2] = fmaxf(m4[2], xv[q].z);
                m4[3] = fmaxf(m4[3], xv[q].w);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                m4[c] = fmaxf(m4[c], __shfl_xor(m4[c], 16, 64));
                m4[c] = fmaxf(m4[c], __shfl_xor(m4[c], 32, 64));
            }
            if (lane < 16) *reinterpret_cast<float4*>(&red[w][4 * xq]) = make_float4(m4[0], m4[1], m4[2], m4[3]);
            __syncthreads();
            float s4[4];
            {
                float4 mx = *reinterpret_cast<const float4*>(&red[0][4 * xq]);
#pragma unroll
                for (int v = 1; v < NW; ++v) {
                    const float4 r = *reinterpret_cast<const float4*>(&red[v][4 * xq]);
                    mx = make_float4(fmaxf(mx.x, r.x), fmaxf(mx.y, r.y), fmaxf(mx.z, r.z), fmaxf(mx.w, r.w));
                }
                s4[0] = neg_shift2(mx.x);
                s4[1] = neg_shift2(mx.y);
                s4[2] = neg_shift2(mx.z);
                s4[3] = neg_shift2(mx.w);
            }
#pragma unroll
            for (int q = 0; q < NQX; ++q) {
                const float x[4] = {xv[q].x, xv[q].y, xv[q].z, xv[q].w};
                T4 ox;
#pragma unroll
                for (int c = 0; c < 4; ++c) ox[c] = MF<T>::cvt(op_exp2(fmaf(x[c], kLog2e, s4[c])));
                *reinterpret_cast<T4*>(&As[rg + RG * q][4 * xq]) = ox;
            }
#pragma unroll
            for (int q = 0; q < NQY; ++q) {
                const float y[4] = {yv[q].x, yv[q].y, yv[q].z, yv[q].w};
                T4 oy;
#pragma unroll
                for (int c = 0; c < 4; ++c) oy[c] = MF<T>::cvt(op_exp2(fmaf(y[c], kLog2e, -s4[c])));
                *reinterpret_cast<T4*>(&Bs[rg + RG * q][4 * xq]) = oy;
            }
            __syncthreads();
            if (b0 + TKB < B) load(b0 + TKB);
#pragma unroll
            for (int kk = 0; kk < TKB; kk += 2 * MF<T>::KPL) {
                const int kc = kk + MF<T>::KPL * (lane >> 5);
                const T* a0 = &As[wm * 64 + (lane & 31)][kc];
                const T* a1 = &As[wm * 64 + 32 + (lane & 31)][kc];
                const T* c0 = &Bs[wn * 64 + (lane & 31)][kc];
                const T* c1 = &Bs[wn * 64 + 32 + (lane & 31)][kc];
                acc[0][0] = MF<T>::mma(a0, c0, acc[0][0]);
                acc[0][1] = MF<T>::mma(a0, c1, acc[0][1]);
                acc[1][0] = MF<T>::mma(a1, c0, acc[1][0]);
                acc[1][1] = MF<T>::mma(a1, c1, acc[1][1]);
            }
        }
    }

    // Epilogue: wavefront w stages its 64 rows x 32 children at a time as [child][row] floats;
    // lane (it, lane) then owns child c = 2 it + (lane >> 5) of the half, rows 2 (lane & 31) + {0, 1}.
    // The quadrant's parameters are loaded first, all at once (padding children read child
    // kt0's, which are not used), so that the epilogue waits for memory once.
    float2 pv[2][16];
#pragma unroll
    for (int sj = 0; sj < 2; ++sj)
#pragma unroll
        for (int it = 0; it < 16; ++it) {
            const int kl = wn * 64 + sj * 32 + 2 * it + (lane >> 5);
            const int64_t p = (int64_t)spid[kl] + m0 + wm * 64 + 2 * (lane & 31);
            pv[sj][it] = make_float2(params[p], params[p + 1]);
        }
    __syncthreads();   // every wavefront is done reading the operand tiles
    float* stg = reinterpret_cast<float*>(ops) + w * 32 * SLD;
#pragma unroll
    for (int sj = 0; sj < 2; ++sj) {
#pragma unroll
        for (int si = 0; si < 2; ++si)
#pragma unroll
            for (int g = 0; g < 4; ++g)
                *reinterpret_cast<float4*>(&stg[(lane & 31) * SLD + si * 32 + 8 * g + 4 * (lane >> 5)]) =
                    make_float4(acc[si][sj][4 * g], acc[si][sj][4 * g + 1], acc[si][sj][4 * g + 2],
                                acc[si][sj][4 * g + 3]);
        __syncthreads();
        // slot offset and product pair of iteration it
        auto item = [&](int it, int64_t& f, float2& v) {
            const int c = 2 * it + (lane >> 5);
            const int kl = wn * 64 + sj * 32 + c;
            const float2 a = *reinterpret_cast<const float2*>(&stg[c * SLD + 2 * (lane & 31)]);
            v = make_float2(sign * pv[sj][it].x * a.x, sign * pv[sj][it].y * a.y);
            f = (int64_t)spf[kl] + m0 + wm * 64 + 2 * (lane & 31);
            return kt0 + kl < n;
        };
        if (mode == kPfOverwrite || mode == kPfAtomic) {
#pragma unroll
            for (int it = 0; it < 16; ++it) {
                int64_t f;
                float2 v;
                if (item(it, f, v)) {
                    pf_add<PF>(&pflows[f], (PF)v.x, mode);
                    pf_add<PF>(&pflows[f + 1], (PF)v.y, mode);
                }
            }
        } else {
            // read-modify-write in two batches of eight iterations (all loads, then all stores)
#pragma unroll
            for (int h = 0; h < 16; h += 8) {
                PF o[8][2];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    int64_t f;
                    float2 v;
                    const bool ok = item(h + u, f, v);
                    o[u][0] = ok ? pflows[f] : (PF)0;
                    o[u][1] = ok ? pflows[f + 1] : (PF)0;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    int64_t f;
                    float2 v;
                    if (item(h + u, f, v)) {
                        pflows[f] = o[u][0] + (PF)v.x;
                        pflows[f + 1] = o[u][1] + (PF)v.y;
                    }
                }
            }
        }
        __syncthreads();
    }
}

// Per-layer launch: logical tile order is m-tile fastest, then child tile, then batch split,
// then node block.
template <typename T, int TM, int TN, typename PF, bool VEC>
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
    par_tile<T, TM, TN, PF, VEC>(pflows, nflows, nmars, emars, params, nids, cids, pids, pfids, nch, K, B, j, m0,
                                 kt0, bstart, bend, modified, atomic_out, sign);
}

// Batched launch over the partitions of many sum layers (one descriptor row per partition, see
// pjamd_sum_backward_par_batched).  The tile order inside a partition is m-tile fastest, then
// child tile, then node block; each tile reduces over the whole batch.  QUAD: the descriptor
// tiles are 128 x 128 and workgroup 4 t + s computes the 64 x 64 quarter s of tile t (the form
// for FP32 operands or B % 4 != 0 under a 128-row table).
template <typename T, int TM, int TN, typename PF, bool VEC, bool QUAD = false>
__global__ __launch_bounds__(256) void sum_par_batched(PF* __restrict__ pflows, const float* __restrict__ nflows,
                                                       const float* __restrict__ nmars,
                                                       const float* __restrict__ emars,
                                                       const float* __restrict__ params, const int* __restrict__ nids,
                                                       const int* __restrict__ cids, const int* __restrict__ pids,
                                                       const int* __restrict__ pfids, const int* __restrict__ nch,
                                                       const int* __restrict__ desc,
                                                       const int* __restrict__ tile_desc, int64_t B, bool modified,
                                                       bool overwrite, float sign) {
    static_assert(!QUAD || (TM == 64 && TN == 64), "quarters of 128 x 128 tiles are 64 x 64");
    const int tq = xcd_tile(blockIdx.x, gridDim.x);
    const int t = QUAD ? tq >> 2 : tq;
    const int* D = desc + (int64_t)tile_desc[t] * PJAMD_PAR_DESC_FIELDS;
    const int local = t - D[0];
    const int K = D[1], mtiles = D[2], ktiles = D[3];
    const int64_t bo = D[4], io = D[5];
    constexpr int SM = QUAD ? 2 * TM : TM, SN = QUAD ? 2 * TN : TN;
    const int m0 = (local % mtiles) * SM + (QUAD ? (tq & 1) * TM : 0);
    const int kt0 = ((local / mtiles) % ktiles) * SN + (QUAD ? ((tq >> 1) & 1) * TN : 0);
    const int j = local / (mtiles * ktiles);
    const int mode = D[7] ? kPfAtomic : (overwrite ? kPfOverwrite : kPfAdd);
    par_tile<T, TM, TN, PF, VEC>(pflows, nflows, nmars, emars + (int64_t)D[6] * B, params, nids + bo, cids + io,
                                 pids + io, pfids + io, nch + bo, K, B, j, m0, kt0, 0, B, modified, mode, sign, D,
                                 D[8]);
}

// Batched launch with 128 x 128 tiles (par_tile_big); same descriptor tables as sum_par_batched.
template <typename T, typename PF>
__global__ __launch_bounds__(BigTile::NT) void sum_par_batched_big(
    PF* __restrict__ pflows, const float* __restrict__ nflows, const float* __restrict__ nmars,
    const float* __restrict__ emars, const float* __restrict__ params, const int* __restrict__ nids,
    const int* __restrict__ cids, const int* __restrict__ pids, const int* __restrict__ pfids,
    const int* __restrict__ nch, const int* __restrict__ desc, const int* __restrict__ tile_desc, int64_t B,
    bool modified, bool overwrite, float sign) {
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int* D = desc + (int64_t)tile_desc[t] * PJAMD_PAR_DESC_FIELDS;
    const int local = t - D[0];
    const int K = D[1], mtiles = D[2], ktiles = D[3];
    const int64_t bo = D[4], io = D[5];
    const int m0 = (local % mtiles) * BigTile::TM;
    const int kt0 = ((local / mtiles) % ktiles) * BigTile::TN;
    const int j = local / (mtiles * ktiles);
    const int mode = D[7] ? kPfAtomic : (overwrite ? kPfOverwrite : kPfAdd);
    par_tile_big<T, PF>(pflows, nflows, nmars, emars + (int64_t)D[6] * B, params, nids + bo, cids + io, pids + io,
                        pfids + io, nch + bo, K, B, j, m0, kt0, modified, mode, sign, D, D[8]);
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
    if (run && g < 256)
        sum_fwd_mfma<T, TM, TN, 2><<<g, 256, 0, stream>>>(node_mars, emars, params, nids, cids, pids, nch, K, B,
                                                          mtiles, btiles);
    else if (run)
        sum_fwd_mfma<T, TM, TN, 1><<<g, 256, 0, stream>>>(node_mars, emars, params, nids, cids, pids, nch, K, B,
                                                          mtiles, btiles);
    else
        sum_fwd_mfma<T, TM, TN, 0><<<g, 256, 0, stream>>>(node_mars, emars, params, nids, cids, pids, nch, K, B,
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
    const int hint = child_run & (PJAMD_LAUNCH_PRE | PJAMD_LAUNCH_NOPRE | PJAMD_LAUNCH_WIDE);
    child_run &= ~(PJAMD_LAUNCH_PRE | PJAMD_LAUNCH_NOPRE | PJAMD_LAUNCH_WIDE);
    const bool aligned = (child_run & PJAMD_PARAMS_ALIGNED) != 0;
    child_run &= ~PJAMD_PARAMS_ALIGNED;
    const bool run = child_run >= TK && child_run % TK == 0;
    bool vec = false;
    if constexpr (sizeof(T) == 2) {
        // 128 x 128 tiles when asked for (measured per layer: fewer bytes from L2 per output, but
        // half the tiles, so fewer loads in flight per CU; faster on some wide layers only)
        const int64_t wide_tiles = (int64_t)n_blocks * (bs / 128) * cdiv(B, 128);
        const bool wide_ok = run && aligned && bs >= 128 && bs % 128 == 0 && B % 4 == 0 && grid_fits(wide_tiles);
        if ((hint & PJAMD_LAUNCH_WIDE) && wide_ok) {
            sum_fwd_big<T><<<(unsigned)wide_tiles, 256, 0, stream>>>(node_mars, emars, params, nids, cids, pids, nch,
                                                                      K, B, bs / 128, (int)cdiv(B, 128));
            PJAMD_LAUNCH_CHECK();
            return PJAMD_OK;
        }
        // node blocks of 32 rows, or child runs shorter than a stage: block-spanning form
        if (aligned && child_run >= 16 && B % 4 == 0 && (bs == 32 || (bs >= 64 && !run))) {
            if (bs == 32)
                sum_fwd_vec_blocks<T, 32, false><<<(unsigned)tiles, 256, 0, stream>>>(
                    node_mars, emars, params, nids, cids, pids, nch, K, B, mtiles, (int)btiles, nullptr, 0);
            else
                sum_fwd_vec_blocks<T, 64, false><<<(unsigned)tiles, 256, 0, stream>>>(
                    node_mars, emars, params, nids, cids, pids, nch, K, B, mtiles, (int)btiles, nullptr, 0);
            PJAMD_LAUNCH_CHECK();
            return PJAMD_OK;
        }
        if (run && aligned && bs >= 64 && B % 4 == 0) {
            vec = true;
            // two stages in flight for grids below the chip's 256 CUs unless the caller says otherwise
            const bool pre = (hint & PJAMD_LAUNCH_PRE) ? true : (hint & PJAMD_LAUNCH_NOPRE) ? false : tiles < 256;
            if (pre)
                sum_fwd_vec<T, true, 0><<<(unsigned)tiles, 256, 0, stream>>>(
                    node_mars, emars, params, nids, cids, pids, nch, K, B, mtiles, (int)btiles, nullptr, nullptr, 0,
                    nullptr);
            else
                sum_fwd_vec<T, false, 0><<<(unsigned)tiles, 256, 0, stream>>>(
                    node_mars, emars, params, nids, cids, pids, nch, K, B, mtiles, (int)btiles, nullptr, nullptr, 0,
                    nullptr);
        }
    }
    if (vec) {
    } else if (bs >= 64)
        launch_fwd_tile<T, 64, 64>(run, (unsigned)tiles, node_mars, emars, params, nids, cids, pids, nch, K, B, mtiles,
                                   (int)btiles, stream);
    else
        launch_fwd_tile<T, 32, 128>(run, (unsigned)tiles, node_mars, emars, params, nids, cids, pids, nch, K, B, mtiles,
                                    (int)btiles, stream);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

template <typename T>
int launch_fwd_split(float* node_mars, const float* emars, const float* params, const int* nids, const int* cids,
                     const int* pids, const int* nch, int n_blocks, int K, int bs, int64_t B, int child_run,
                     int nsplit, float* ws, hipStream_t stream) {
    const bool aligned = (child_run & PJAMD_PARAMS_ALIGNED) != 0;
    child_run &= ~PJAMD_PARAMS_ALIGNED;
    if (!(aligned && child_run >= TK && child_run % TK == 0 && bs >= 64 && bs % 64 == 0 && B % 4 == 0))
        return PJAMD_EUNSUPPORTED;
    const int mtiles = bs / 64;
    const int64_t btiles = cdiv(B, 64);
    const int64_t tiles = (int64_t)n_blocks * mtiles * btiles;
    PJAMD_REQUIRE(grid_fits(tiles * nsplit));
    float* ws_acc = ws;
    float* ws_max = ws + (int64_t)nsplit * n_blocks * bs * B;
    sum_fwd_vec_split<T><<<(unsigned)(tiles * nsplit), 256, 0, stream>>>(node_mars, emars, params, nids, cids, pids,
                                                                          nch, K, B, mtiles, (int)btiles, nsplit,
                                                                          ws_acc, ws_max);
    PJAMD_LAUNCH_CHECK();
    const int64_t rows = (int64_t)n_blocks * bs;
    PJAMD_REQUIRE(rows < 65536);
    PJAMD_REQUIRE(nsplit <= kMaxSplits);
    sum_fwd_split_combine<<<(unsigned)cdiv(rows * (B / 4), 256), 256, 0, stream>>>(node_mars, nids, ws_acc, ws_max,
                                                                                   nsplit, bs, rows, B);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

template <typename T>
int launch_ele_split(float* node_flows, const float* nmars, const float* emars, const float* params,
                     const int* chids, const int* parids, const int* parpids, const int* npar, const int* ch_nodes,
                     int Kc, int n, int KP, int cbs, int bs, int64_t B, bool mod, bool modc, bool aligned, int nsplit,
                     float* ws, hipStream_t stream) {
    const int lbs = log2_exact(bs);
    if (!(aligned && lbs >= 0 && bs >= TK && cbs >= 64 && cbs % 64 == 0 && B % 4 == 0)) return PJAMD_EUNSUPPORTED;
    const int mtiles = cbs / 64;
    const int64_t btiles = cdiv(B, 64);
    const int64_t tiles = (int64_t)n * mtiles * btiles;
    const int64_t rows = (int64_t)n * cbs;
    PJAMD_REQUIRE(grid_fits(tiles * nsplit) && rows < 65536);
    float* ws_acc = ws;
    float* ws_max = ws + (int64_t)nsplit * rows * B;
    sum_ele_vec_split<T><<<(unsigned)(tiles * nsplit), 256, 0, stream>>>(node_flows, nmars, params, parids, parpids,
                                                                          npar, KP, lbs, B, mtiles, (int)btiles, mod,
                                                                          nsplit, ws_acc, ws_max);
    PJAMD_LAUNCH_CHECK();
    PJAMD_REQUIRE(nsplit <= kMaxSplits);
    sum_ele_split_combine<<<(unsigned)cdiv(rows * (B / 4), 256), 256, 0, stream>>>(
        node_flows, nmars, emars, chids, ch_nodes, Kc, modc, ws_acc, ws_max, nsplit, cbs, rows, B);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

// Sum forward with the product layer below fused in (sum_fwd_vec with KP > 0).  Only the 16-byte
// RUN form exists: the caller checks the layout (runs of >= TK children inside one product block)
// and calls the plain product and sum forwards otherwise.
template <typename T, int KP>
void launch_fwd_prod_kp(bool pre, unsigned tiles, float* node_mars, float* emars, const float* params,
                        const int* nids, const int* cids, const int* pids, const int* nch, int K, int64_t B,
                        int mtiles, int btiles, const int* pch, int nstages, hipStream_t stream) {
    if (pre)
        sum_fwd_vec<T, true, KP><<<tiles, 256, 0, stream>>>(node_mars, emars, params, nids, cids, pids, nch, K, B,
                                                            mtiles, btiles, node_mars, pch, nstages, emars);
    else
        sum_fwd_vec<T, false, KP><<<tiles, 256, 0, stream>>>(node_mars, emars, params, nids, cids, pids, nch, K, B,
                                                             mtiles, btiles, node_mars, pch, nstages, emars);
}

template <typename T>
int launch_fwd_prod(float* node_mars, float* emars, const float* params, const int* nids, const int* cids,
                    const int* pids, const int* nch, int n_blocks, int K, int bs, int64_t B, int child_run,
                    const int* pch, int kp, hipStream_t stream) {
    const bool aligned = (child_run & PJAMD_PARAMS_ALIGNED) != 0;
    child_run &= ~PJAMD_PARAMS_ALIGNED;
    if (!(aligned && child_run >= TK && child_run % TK == 0 && bs >= 64 && bs % 64 == 0 && B % 4 == 0))
        return PJAMD_EUNSUPPORTED;
    const int mtiles = bs / 64;
    const int64_t btiles = cdiv(B, 64);
    const int64_t tiles = (int64_t)n_blocks * mtiles * btiles;
    PJAMD_REQUIRE(grid_fits(tiles));
    const int nstages = (int)cdiv(K, TK);
    const bool pre = tiles < 256;
    switch (kp) {
        case 2: launch_fwd_prod_kp<T, 2>(pre, (unsigned)tiles, node_mars, emars, params, nids, cids, pids, nch, K, B,
                                         mtiles, (int)btiles, pch, nstages, stream); break;
        case 4: launch_fwd_prod_kp<T, 4>(pre, (unsigned)tiles, node_mars, emars, params, nids, cids, pids, nch, K, B,
                                         mtiles, (int)btiles, pch, nstages, stream); break;
        case 8: launch_fwd_prod_kp<T, 8>(pre, (unsigned)tiles, node_mars, emars, params, nids, cids, pids, nch, K, B,
                                         mtiles, (int)btiles, pch, nstages, stream); break;
        default: return PJAMD_EUNSUPPORTED;
    }
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

template <typename T, int TM, int TN, int FUSE>
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
               bool aligned, hipStream_t stream, int hint = 0) {
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
    if constexpr (sizeof(T) == 2) {
        // child blocks of 32 rows, or parent blocks shorter than a stage: block-spanning form
        if (aligned && mod && B % 4 == 0 && bs >= 16 && KP <= kEleIdx && (cbs == 32 || (cbs >= 64 && !ob))) {
#define PJAMD_ELE_BLK(TM_, F_)                                                                                    \
    sum_ele_vec_blocks<T, TM_, F_><<<g, 256, 0, stream>>>(eflows, nflows, nmars, emars, params, chids, parids,     \
                                                          parpids, npar, KP, lbs, B, mtiles, bt, accum, child_flows, \
                                                          ch_nodes, Kc, mod_children)
            if (cbs == 32) {
                if (ch_nodes != nullptr) PJAMD_ELE_BLK(32, 1);
                else PJAMD_ELE_BLK(32, 0);
            } else {
                if (ch_nodes != nullptr) PJAMD_ELE_BLK(64, 1);
                else PJAMD_ELE_BLK(64, 0);
            }
#undef PJAMD_ELE_BLK
            PJAMD_LAUNCH_CHECK();
            return PJAMD_OK;
        }
        if (aligned && ob && cbs >= 64 && B % 4 == 0) {
#define PJAMD_ELE_VEC(F_, P_)                                                                                     \
    sum_ele_vec<T, F_, P_><<<g, 256, 0, stream>>>(eflows, nflows, nmars, emars, params, chids, parids, parpids, npar, \
                                                  KP, lbs, B, mtiles, bt, mod, accum, child_flows, ch_nodes, Kc,  \
                                                  mod_children)
            const bool pre = (hint & PJAMD_LAUNCH_PRE) ? true : (hint & PJAMD_LAUNCH_NOPRE) ? false : g < 256;
            if (ch_nodes != nullptr) {
                if (pre && Kc <= kPreChildren) PJAMD_ELE_VEC(2, true);
                else if (pre) PJAMD_ELE_VEC(1, true);
                else PJAMD_ELE_VEC(1, false);
            } else {
                if (pre) PJAMD_ELE_VEC(0, true);
                else PJAMD_ELE_VEC(0, false);
            }
#undef PJAMD_ELE_VEC
            PJAMD_LAUNCH_CHECK();
            return PJAMD_OK;
        }
    }
    // 0: element flows; 1: children scatter; 2: children scatter with prefetched children (grids
    // smaller than the chip's 256 CUs, where every dependent round trip is paid in full)
    const int fuse = ch_nodes == nullptr ? 0 : ((Kc <= kPreChildren && g < 256 && ob) ? 2 : 1);
#define PJAMD_ELE_TILE(TM_, TN_, F_)                                                                              \
    launch_ele_tile<T, TM_, TN_, F_>(ob, g, eflows, nflows, nmars, emars, params, chids, parids, parpids, npar, \
                                     KP, lbs, B, mtiles, bt, mod, accum, child_flows, ch_nodes, Kc, mod_children, \
                                     stream)
    if (cbs >= 64) {
        if (fuse == 2) PJAMD_ELE_TILE(64, 64, 2);
        else if (fuse == 1) PJAMD_ELE_TILE(64, 64, 1);
        else PJAMD_ELE_TILE(64, 64, 0);
    } else {
        if (fuse == 2) PJAMD_ELE_TILE(32, 128, 2);
        else if (fuse == 1) PJAMD_ELE_TILE(32, 128, 1);
        else PJAMD_ELE_TILE(32, 128, 0);
    }
#undef PJAMD_ELE_TILE
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
#define PJAMD_PAR(TM_, TN_, V_)                                                                                   \
    sum_par_mfma<T, TM_, TN_, PF, V_><<<grid, 256, 0, stream>>>(pflows, nflows, nmars, emars, params, nids, cids, pids, \
                                                               pfids, nch, K, B, bchunk, mtiles, ktiles, (int)splits, \
                                                               mod, atomic_out, sign)
    // 16-byte loads along the batch need B % 4 == 0 (split bounds are multiples of TK)
    const bool vec = sizeof(T) == 2 && B % 4 == 0;
    if (bs >= 64) {
        if (vec) PJAMD_PAR(64, 64, sizeof(T) == 2);
        else PJAMD_PAR(64, 64, false);
    } else {
        if (vec) PJAMD_PAR(32, 128, sizeof(T) == 2);
        else PJAMD_PAR(32, 128, false);
    }
#undef PJAMD_PAR
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

template <typename T, typename PF>
int launch_par_batched(PF* pflows, const float* nflows, const float* nmars, const float* emars, const float* params,
                       const int* nids, const int* cids, const int* pids, const int* pfids, const int* nch,
                       const int* desc, const int* tile_desc, int n_tiles, int tile_rows, int64_t B, bool mod,
                       bool overwrite, float sign, hipStream_t stream) {
    PJAMD_REQUIRE(grid_fits(n_tiles));
#define PJAMD_PARB(TM_, TN_, V_)                                                                                  \
    sum_par_batched<T, TM_, TN_, PF, V_><<<(unsigned)n_tiles, 256, 0, stream>>>(                                  \
        pflows, nflows, nmars, emars, params, nids, cids, pids, pfids, nch, desc, tile_desc, B, mod, overwrite, sign)
    const bool vec = sizeof(T) == 2 && B % 4 == 0;
    if (tile_rows == 128) {
        if constexpr (sizeof(T) == 2) {
            if (vec) {
                sum_par_batched_big<T, PF><<<(unsigned)n_tiles, BigTile::NT, 0, stream>>>(
                    pflows, nflows, nmars, emars, params, nids, cids, pids, pfids, nch, desc, tile_desc, B, mod,
                    overwrite, sign);
                PJAMD_LAUNCH_CHECK();
                return PJAMD_OK;
            }
        }
        PJAMD_REQUIRE(grid_fits(4 * (int64_t)n_tiles));
        sum_par_batched<T, 64, 64, PF, false, true><<<(unsigned)(4 * n_tiles), 256, 0, stream>>>(
            pflows, nflows, nmars, emars, params, nids, cids, pids, pfids, nch, desc, tile_desc, B, mod, overwrite,
            sign);
    } else if (tile_rows == 64) {
        if (vec) PJAMD_PARB(64, 64, sizeof(T) == 2);
        else PJAMD_PARB(64, 64, false);
    } else {
        if (vec) PJAMD_PARB(32, 128, sizeof(T) == 2);
        else PJAMD_PARB(32, 128, false);
    }
#undef PJAMD_PARB
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

PJAMD_EXPORT int pjamd_sum_forward_split(float* node_mars, const float* element_mars, const float* params,
                                         const int* nids, const int* cids, const int* pids, const int* n_chs,
                                         int n_blocks, int K, int bs, int64_t B, int child_run, int n_splits,
                                         float* workspace, int precision, hipStream_t stream) {
    PJAMD_REQUIRE(n_blocks >= 0 && K >= 1 && bs >= 1 && B > 0 && n_splits >= 2 && n_splits <= 16);
    if (n_blocks == 0) return PJAMD_OK;
    if (precision == 1)
        return launch_fwd_split<__bf16>(node_mars, element_mars, params, nids, cids, pids, n_chs, n_blocks, K, bs, B,
                                        child_run, n_splits, workspace, stream);
    if (precision == 2)
        return launch_fwd_split<_Float16>(node_mars, element_mars, params, nids, cids, pids, n_chs, n_blocks, K, bs,
                                          B, child_run, n_splits, workspace, stream);
    return PJAMD_EUNSUPPORTED;
}

PJAMD_EXPORT int pjamd_sum_backward_ele_split(float* node_flows, const float* node_mars, const float* element_mars,
                                              const float* params, const int* chids, const int* parids,
                                              const int* parpids, const int* n_pars, const int* ch_nodes, int Kc,
                                              int n_cblocks, int KP, int cbs, int bs, int64_t B, int flags,
                                              int n_splits, float* workspace, int precision, hipStream_t stream) {
    PJAMD_REQUIRE(n_cblocks >= 0 && KP >= 1 && Kc >= 1 && bs >= 1 && B > 0 && ch_nodes != nullptr && n_splits >= 2 &&
                  n_splits <= 16);
    if (n_cblocks == 0) return PJAMD_OK;
    const bool mod = flags & 1, modc = (flags & 2) != 0, aligned = (flags & 4) != 0;
    if (precision == 1)
        return launch_ele_split<__bf16>(node_flows, node_mars, element_mars, params, chids, parids, parpids, n_pars,
                                        ch_nodes, Kc, n_cblocks, KP, cbs, bs, B, mod, modc, aligned, n_splits,
                                        workspace, stream);
    if (precision == 2)
        return launch_ele_split<_Float16>(node_flows, node_mars, element_mars, params, chids, parids, parpids, n_pars,
                                          ch_nodes, Kc, n_cblocks, KP, cbs, bs, B, mod, modc, aligned, n_splits,
                                          workspace, stream);
    return PJAMD_EUNSUPPORTED;
}

PJAMD_EXPORT int pjamd_sum_forward_prod(float* node_mars, float* element_mars, const float* params, const int* nids,
                                        const int* cids, const int* pids, const int* n_chs, int n_blocks, int K,
                                        int bs, int64_t B, int child_run, const int* prod_bases, int prod_k,
                                        int precision, hipStream_t stream) {
    PJAMD_REQUIRE(n_blocks >= 0 && K >= 1 && bs >= 1 && B > 0 && prod_k >= 1);
    if (n_blocks == 0) return PJAMD_OK;
    if (precision == 1)
        return launch_fwd_prod<__bf16>(node_mars, element_mars, params, nids, cids, pids, n_chs, n_blocks, K, bs, B,
                                       child_run, prod_bases, prod_k, stream);
    if (precision == 2)
        return launch_fwd_prod<_Float16>(node_mars, element_mars, params, nids, cids, pids, n_chs, n_blocks, K, bs,
                                         B, child_run, prod_bases, prod_k, stream);
    return PJAMD_EUNSUPPORTED;
}

int pjamd_sum_backward_ele_mfma(float* eflows, const float* nflows, const float* nmars, const float* emars,
                                const float* params, const int* chids, const int* parids, const int* parpids,
                                const int* npar, int n, int KP, int cbs, int bs, int64_t B, bool mod, bool accum,
                                bool aligned, int precision, hipStream_t stream) {
    if (cbs < 32) return PJAMD_EUNSUPPORTED;
    if (precision == 1)
        return launch_ele<__bf16>(eflows, nflows, nmars, emars, params, chids, parids, parpids, npar, n, KP, cbs, bs, B,
                                  mod, accum, nullptr, nullptr, 0, false, aligned, stream);
    if (precision == 2)
        return launch_ele<_Float16>(eflows, nflows, nmars, emars, params, chids, parids, parpids, npar, n, KP, cbs, bs,
                                    B, mod, accum, nullptr, nullptr, 0, false, aligned, stream);
    return launch_ele<float>(eflows, nflows, nmars, emars, params, chids, parids, parpids, npar, n, KP, cbs, bs, B,
                             mod, accum, nullptr, nullptr, 0, false, aligned, stream);
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
    const bool mod = flags & 1, modc = (flags & 2) != 0, aligned = (flags & 4) != 0;
    const int hint = flags & (PJAMD_LAUNCH_PRE | PJAMD_LAUNCH_NOPRE);
    if (precision == 1)
        return launch_ele<__bf16>(nullptr, node_flows, node_mars, element_mars, params, chids, parids, parpids, n_pars,
                                  n_cblocks, KP, cbs, bs, B, mod, false, node_flows, ch_nodes, Kc, modc, aligned, stream,
                                  hint);
    if (precision == 2)
        return launch_ele<_Float16>(nullptr, node_flows, node_mars, element_mars, params, chids, parids, parpids,
                                    n_pars, n_cblocks, KP, cbs, bs, B, mod, false, node_flows, ch_nodes, Kc, modc,
                                    aligned, stream, hint);
    return launch_ele<float>(nullptr, node_flows, node_mars, element_mars, params, chids, parids, parpids, n_pars,
                             n_cblocks, KP, cbs, bs, B, mod, false, node_flows, ch_nodes, Kc, modc, aligned, stream,
                             hint);
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
    PJAMD_REQUIRE(n_tiles >= 0 && B > 0 && (tile_rows == 32 || tile_rows == 64 || tile_rows == 128) &&
                  precision >= 0 &&
                  precision <= 2);
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

PJAMD_EXPORT int pjamd_sum_forward_queue(float* node_mars, const float* element_mars, const float* params,
                                        const int* nids, const int* cids, const int* pids, const int* n_chs,
                                        int n_blocks, int K, int bs, int64_t B, int child_run, int* queue,
                                        int precision, hipStream_t stream) {
    PJAMD_REQUIRE(n_blocks >= 0 && K >= 1 && bs >= 1 && B > 0 && queue != nullptr && child_run >= 0);
    if (n_blocks == 0) return PJAMD_OK;
    const bool aligned = (child_run & PJAMD_PARAMS_ALIGNED) != 0;
    const int run = child_run & ~(PJAMD_PARAMS_ALIGNED | PJAMD_LAUNCH_PRE | PJAMD_LAUNCH_NOPRE);
    if (!((precision == 1 || precision == 2) && aligned && run >= 16 && run % 16 == 0 && B % 4 == 0 &&
          (bs == 32 || (bs >= 64 && bs % 64 == 0))))
        return PJAMD_EUNSUPPORTED;
    const int TM = bs == 32 ? 32 : 64, TN = 4096 / TM;
    const int mtiles = bs / TM;
    const int64_t btiles = cdiv(B, TN);
    const int64_t tiles = (int64_t)n_blocks * mtiles * btiles;
    PJAMD_REQUIRE(grid_fits(tiles));
    // as many workgroups as the chip keeps resident (occupancy of the kernel times the CUs)
    static int resident[2][2] = {{0, 0}, {0, 0}};
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return (int)e;
    int& occ = resident[precision - 1][TM == 32 ? 0 : 1];
    if (occ == 0) {
        int o = 0;
        if (precision == 1)
            e = TM == 32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, sum_fwd_vec_blocks<__bf16, 32, true>, 256, 0)
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, sum_fwd_vec_blocks<__bf16, 64, true>, 256, 0);
        else
            e = TM == 32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, sum_fwd_vec_blocks<_Float16, 32, true>, 256, 0)
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, sum_fwd_vec_blocks<_Float16, 64, true>, 256, 0);
        if (e != hipSuccess) return (int)e;
        occ = o > 0 ? o : 1;
    }
    const int64_t workers = (int64_t)occ * cus;
    const unsigned grid = (unsigned)(tiles < workers ? tiles : workers);
#define PJAMD_FWD_Q(T_, TM_)                                                                                         \
    sum_fwd_vec_blocks<T_, TM_, true><<<grid, 256, 0, stream>>>(node_mars, element_mars, params, nids, cids, pids, \
                                                                n_chs, K, B, mtiles, (int)btiles, queue, (int)tiles)
    if (precision == 1) {
        if (TM == 32) PJAMD_FWD_Q(__bf16, 32);
        else PJAMD_FWD_Q(__bf16, 64);
    } else {
        if (TM == 32) PJAMD_FWD_Q(_Float16, 32);
        else PJAMD_FWD_Q(_Float16, 64);
    }
#undef PJAMD_FWD_Q
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

// ---------------------------------------------------------------------------------------------
// Allocation-free split forward.  Its workspace is `scratch`, a fixed device-resident array of
// PJAMD_SPLIT_SCRATCH_FLOATS floats (32 MiB) that every device the library runs on holds until the
// process ends.  In host code `scratch` names only a host-side stand-in: its device address is
// looked up with hipGetSymbolAddress once per device (with that device current) and kept in
// g_scratch_ptr; only that pointer is ever handed to a kernel.  The split kernel writes every
// partial and column maximum the combine reads, so nothing needs clearing between calls; calls
// that overlap on one device share the array and race.
static __device__ __align__(256) float scratch[PJAMD_SPLIT_SCRATCH_FLOATS];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device.
int current_scratch(float** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kMaxScratchDevices) return PJAMD_EINVAL;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return PJAMD_OK;
}
}  // namespace

PJAMD_EXPORT int pjamd_sum_forward_split_noalloc(float* node_mars, const float* element_mars, const float* params,
                                                 const int* nids, const int* cids, const int* pids, const int* n_chs,
                                                 int n_blocks, int K, int bs, int64_t B, int child_run, int n_splits,
                                                 int precision, hipStream_t stream) {
    PJAMD_REQUIRE(n_blocks >= 0 && K >= 1 && bs >= 1 && B > 0 && n_splits >= 2 && n_splits <= 16);
    if (n_blocks == 0) return PJAMD_OK;
    const int64_t rows = (int64_t)n_blocks * bs;
    const int64_t need = (int64_t)n_splits * (rows + (rows + 63) / 64) * B;
    if (need > (int64_t)PJAMD_SPLIT_SCRATCH_FLOATS) return PJAMD_ETOOLARGE;
    float* ws = nullptr;
    const int st = current_scratch(&ws);
    if (st != PJAMD_OK) return st;
    return pjamd_sum_forward_split(node_mars, element_mars, params, nids, cids, pids, n_chs, n_blocks, K, bs, B,
                                   child_run, n_splits, ws, precision, stream);
}
