// Sum layers in FP32: log-sum-exp forward, in-place flow modification, element-flow and
// parameter-flow backward.
//
// Layout: node / element rows are [*, B] (batch contiguous).  A sum node block j has `bs`
// nodes starting at nids[j]; its child k (k < nch[j], padding above) is element row cids[j,k]
// with the `bs` parameters theta[pids[j,k] + i] (parent index fastest).
//
// All three contractions are LDS-tiled GEMMs with per-column max subtraction:
//   forward   M = parent nodes, N = batch,    K = children   (running max, rescaled accumulator)
//   ele-flows M = child nodes,  N = batch,    K = parents    (running max, rescaled accumulator)
//   par-flows M = parent nodes, N = children, K = batch      (per-chunk shift, exact product)
#include "pjamd_common.h"

using namespace pjamd;

namespace {

constexpr int kThreads = 256;
constexpr int kTN = 64;   // batch columns per workgroup (forward / ele-flows)
constexpr int kTK = 32;   // reduction depth per LDS stage

// ---------------------------------------------------------------------------------------------
// Forward, bs >= 16.  Workgroup = (node block j, TM parent rows) x 64 batch columns.
template <int TM>
__global__ __launch_bounds__(kThreads) void sum_fwd_tiled(float* __restrict__ node_mars, const float* __restrict__ emars,
                                                          const float* __restrict__ params, const int* __restrict__ nids,
                                                          const int* __restrict__ cids, const int* __restrict__ pids,
                                                          const int* __restrict__ nch, int K, int bs, int64_t B) {
    constexpr int RM = TM / 16, RN = kTN / 16;
    __shared__ float As[kTK][TM + 4];
    __shared__ float Bs[kTK][kTN + 4];
    __shared__ float red[kThreads / kTN][kTN];
    __shared__ float mrun[kTN], mscale[kTN];

    const int mtiles = bs / TM;
    const int j = blockIdx.x / mtiles;
    const int m0 = (blockIdx.x % mtiles) * TM;
    const int64_t b0 = (int64_t)blockIdx.y * kTN;
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int n = nch[j];
    const int* crow = cids + (int64_t)j * K;
    const int* prow = pids + (int64_t)j * K;

    // loader mappings
    const int lcol = tid % kTN, lrg = tid / kTN;         // element tile: 4 row groups x 64 columns
    const int64_t lb = b0 + lcol;
    const int acol = tid % TM, ark = tid / TM;           // parameter tile: (256 / TM) k groups x TM rows
    constexpr int AKG = kThreads / TM;

    if (tid < kTN) mrun[tid] = kNegInf;
    float acc[RM][RN];
#pragma unroll
    for (int a = 0; a < RM; ++a)
#pragma unroll
        for (int c = 0; c < RN; ++c) acc[a][c] = 0.0f;

    for (int k0 = 0; k0 < n; k0 += kTK) {
        float ev[kTK / 4];
        float lmax = kNegInf;
#pragma unroll
        for (int r = 0; r < kTK / 4; ++r) {
            const int kk = lrg + r * 4;
            const int k = k0 + kk;
            float v = kNegInf;
            if (k < n && lb < B) v = emars[(int64_t)crow[k] * B + lb];
            ev[r] = v;
            lmax = fmaxf(lmax, v);
        }
        red[lrg][lcol] = lmax;
        __syncthreads();
        if (tid < kTN) {
            float t = fmaxf(fmaxf(red[0][tid], red[1][tid]), fmaxf(red[2][tid], red[3][tid]));
            const float mo = mrun[tid];
            const float mn = fmaxf(mo, t);
            mscale[tid] = (mn == kNegInf) ? 1.0f : __expf(mo - mn);
            mrun[tid] = mn;
        }
        __syncthreads();
        {
            const float mn = mrun[lcol];
#pragma unroll
            for (int r = 0; r < kTK / 4; ++r) {
                const int kk = lrg + r * 4;
                Bs[kk][lcol] = (mn == kNegInf) ? 0.0f : __expf(ev[r] - mn);
            }
        }
#pragma unroll
        for (int r = 0; r < kTK / AKG; ++r) {
            const int kk = ark + r * AKG;
            const int k = k0 + kk;
            As[kk][acol] = (k < n) ? params[(int64_t)prow[k] + m0 + acol] : 0.0f;
        }
#pragma unroll
        for (int c = 0; c < RN; ++c) {
            const float s = mscale[tx + 16 * c];
#pragma unroll
            for (int a = 0; a < RM; ++a) acc[a][c] *= s;
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < kTK; ++kk) {
            float av[RM], bv[RN];
#pragma unroll
            for (int a = 0; a < RM; ++a) av[a] = As[kk][ty + 16 * a];
#pragma unroll
            for (int c = 0; c < RN; ++c) bv[c] = Bs[kk][tx + 16 * c];
#pragma unroll
            for (int a = 0; a < RM; ++a)
#pragma unroll
                for (int c = 0; c < RN; ++c) acc[a][c] = fmaf(av[a], bv[c], acc[a][c]);
        }
        __syncthreads();
    }
    const int64_t nrow0 = (int64_t)nids[j] + m0;
#pragma unroll
    for (int c = 0; c < RN; ++c) {
        const int col = tx + 16 * c;
        const int64_t b = b0 + col;
        if (b >= B) continue;
        const float m = mrun[col];
#pragma unroll
        for (int a = 0; a < RM; ++a) {
            const float v = acc[a][c];
            node_mars[(nrow0 + ty + 16 * a) * B + b] = (v > 0.0f && m != kNegInf) ? __logf(v) + m : kNegInf;
        }
    }
}

// Forward, any bs (used for bs < 16): one workgroup per (node, 64 batch columns); its W
// wavefronts split the children in chunks of kRowChunk (wavefront w takes chunks w, w + W, ...),
// lanes run along the batch.  A chunk's index, element and parameter loads are all issued before
// the first use (indices clamped to the last real child, extra entries masked afterwards), and the
// chunk is folded into a running (max, sum of theta * exp(e - max)) pair, so a node with hundreds
// of children (a circuit's root) costs a few memory round trips instead of one per child.  The
// wavefronts' pairs are combined through LDS.
constexpr int kRowChunk = 16;
constexpr int kSmallWaves = kThreads / 64;   // wavefronts of the other row-wise kernels

template <int W>
__global__ __launch_bounds__(64 * W) void sum_fwd_rowwise(float* __restrict__ node_mars, const float* __restrict__ emars,
                                                          const float* __restrict__ params, const int* __restrict__ nids,
                                                          const int* __restrict__ cids, const int* __restrict__ pids,
                                                          const int* __restrict__ nch, int K, int bs, int64_t B) {
    __shared__ float rm[W][64], rs[W][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t b = (int64_t)blockIdx.y * 64 + lane;
    const bool ok = b < B;
    const int64_t bc = ok ? b : 0;      // in-bounds column for lanes past the batch (result unused)
    const int j = blockIdx.x / bs, i = blockIdx.x % bs;
    const int n = nch[j];
    const int* crow = cids + (int64_t)j * K;
    const int* prow = pids + (int64_t)j * K;
    float m = kNegInf, s = 0.0f;
    for (int k0 = w * kRowChunk; k0 < n; k0 += W * kRowChunk) {
        float e[kRowChunk], p[kRowChunk];
#pragma unroll
        for (int q = 0; q < kRowChunk; ++q) {
            const int k = min(k0 + q, n - 1);
            e[q] = emars[(int64_t)crow[k] * B + bc];
            p[q] = params[(int64_t)prow[k] + i];
        }
        float mc = kNegInf;
#pragma unroll
        for (int q = 0; q < kRowChunk; ++q) mc = (k0 + q < n) ? fmaxf(mc, e[q]) : mc;
        const float mn = fmaxf(m, mc);
        if (mn != kNegInf) {
            s *= __expf(m - mn);
#pragma unroll
            for (int q = 0; q < kRowChunk; ++q) s = (k0 + q < n) ? fmaf(p[q], __expf(e[q] - mn), s) : s;
            m = mn;
        }
    }
    rm[w][lane] = m;
    rs[w][lane] = s;
    __syncthreads();
    if (w == 0 && ok) {
        float M = kNegInf;
#pragma unroll
        for (int q = 0; q < W; ++q) M = fmaxf(M, rm[q][lane]);
        float S = 0.0f;
        if (M != kNegInf) {
#pragma unroll
            for (int q = 0; q < W; ++q) S += (rm[q][lane] == kNegInf) ? 0.0f : rs[q][lane] * __expf(rm[q][lane] - M);
        }
        node_mars[((int64_t)nids[j] + i) * B + b] = (S > 0.0f && M != kNegInf) ? __logf(S) + M : kNegInf;
    }
}

// node_flows <- log(node_flows) - node_mars on the layer's nodes (-inf where the marginal is -inf).
__global__ __launch_bounds__(kThreads) void modify_flows_kernel(float* __restrict__ nflows, const float* __restrict__ nmars,
                                                                const int* __restrict__ nids, int bs, int64_t B) {
    const int64_t b = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int j = blockIdx.x / bs, i = blockIdx.x % bs;
    const int64_t o = ((int64_t)nids[j] + i) * B + b;
    const float m = nmars[o];
    nflows[o] = (m == kNegInf) ? kNegInf : __logf(nflows[o]) - m;
}

__device__ __forceinline__ float log_flow_div_mar(const float* nflows, const float* nmars, int64_t o, bool modified) {
    if (modified) return nflows[o];
    const float m = nmars[o];
    return (m == kNegInf) ? kNegInf : __logf(nflows[o]) - m;
}

// ---------------------------------------------------------------------------------------------
// Element flows: ef[c0 + ci, b] = exp(e[c0 + ci, b]) * sum_{k, pi} theta[parpid_k + ci*bs + pi]
//                                 * exp(lfdm[parid_k + pi, b]).
// Workgroup = (child block j, TM child rows) x 64 batch columns; the reduction runs over the
// flattened (parent block k, parent row pi) index.
template <int TM>
__global__ __launch_bounds__(kThreads) void sum_bwd_ele_tiled(float* __restrict__ eflows, const float* __restrict__ nflows,
                                                              const float* __restrict__ nmars, const float* __restrict__ emars,
                                                              const float* __restrict__ params, const int* __restrict__ chids,
                                                              const int* __restrict__ parids, const int* __restrict__ parpids,
                                                              const int* __restrict__ npar, int KP, int cbs, int bs,
                                                              int64_t B, bool modified, bool accumulate) {
    constexpr int RM = TM / 16, RN = kTN / 16;
    __shared__ float As[kTK][TM + 4];
    __shared__ float Bs[kTK][kTN + 4];
    __shared__ float red[kThreads / kTN][kTN];
    __shared__ float mrun[kTN], mscale[kTN];

    const int mtiles = cbs / TM;
    const int j = blockIdx.x / mtiles;
    const int m0 = (blockIdx.x % mtiles) * TM;
    const int64_t b0 = (int64_t)blockIdx.y * kTN;
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int kflat = npar[j] * bs;
    const int* parrow = parids + (int64_t)j * KP;
    const int* pprow = parpids + (int64_t)j * KP;
    const int lcol = tid % kTN, lrg = tid / kTN;
    const int64_t lb = b0 + lcol;
    // parameter loader: rows ci (stride bs in memory), consecutive kk within a row
    const int akk = tid % kTK, aci = tid / kTK;   // 32 kk x 8 rows per pass
    constexpr int APASS = kThreads / kTK;

    if (tid < kTN) mrun[tid] = kNegInf;
    float acc[RM][RN];
#pragma unroll
    for (int a = 0; a < RM; ++a)
#pragma unroll
        for (int c = 0; c < RN; ++c) acc[a][c] = 0.0f;

    for (int k0 = 0; k0 < kflat; k0 += kTK) {
        float lv[kTK / 4];
        float lmax = kNegInf;
#pragma unroll
        for (int r = 0; r < kTK / 4; ++r) {
            const int kk = k0 + lrg + r * 4;
            float v = kNegInf;
            if (kk < kflat && lb < B) {
                const int kb = kk / bs, pi = kk - kb * bs;
                v = log_flow_div_mar(nflows, nmars, ((int64_t)parrow[kb] + pi) * B + lb, modified);
            }
            lv[r] = v;
            lmax = fmaxf(lmax, v);
        }
        red[lrg][lcol] = lmax;
        __syncthreads();
        if (tid < kTN) {
            float t = fmaxf(fmaxf(red[0][tid], red[1][tid]), fmaxf(red[2][tid], red[3][tid]));
            const float mo = mrun[tid];
            const float mn = fmaxf(mo, t);
            mscale[tid] = (mn == kNegInf) ? 1.0f : __expf(mo - mn);
            mrun[tid] = mn;
        }
        __syncthreads();
        {
            const float mn = mrun[lcol];
#pragma unroll
            for (int r = 0; r < kTK / 4; ++r) Bs[lrg + r * 4][lcol] = (mn == kNegInf) ? 0.0f : __expf(lv[r] - mn);
        }
        {
            const int kk = k0 + akk;
            int kb = 0, pi = 0;
            const bool ok = kk < kflat;
            if (ok) { kb = kk / bs; pi = kk - kb * bs; }
            const int64_t pbase = ok ? (int64_t)pprow[kb] + pi : 0;
#pragma unroll
            for (int r = 0; r < TM / APASS; ++r) {
                const int ci = aci + r * APASS;
                As[akk][ci] = ok ? params[pbase + (int64_t)(m0 + ci) * bs] : 0.0f;
            }
        }
#pragma unroll
        for (int c = 0; c < RN; ++c) {
            const float s = mscale[tx + 16 * c];
#pragma unroll
            for (int a = 0; a < RM; ++a) acc[a][c] *= s;
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < kTK; ++kk) {
            float av[RM], bv[RN];
#pragma unroll
            for (int a = 0; a < RM; ++a) av[a] = As[kk][ty + 16 * a];
#pragma unroll
            for (int c = 0; c < RN; ++c) bv[c] = Bs[kk][tx + 16 * c];
#pragma unroll
            for (int a = 0; a < RM; ++a)
#pragma unroll
                for (int c = 0; c < RN; ++c) acc[a][c] = fmaf(av[a], bv[c], acc[a][c]);
        }
        __syncthreads();
    }
    const int64_t crow0 = (int64_t)chids[j] + m0;
#pragma unroll
    for (int c = 0; c < RN; ++c) {
        const int col = tx + 16 * c;
        const int64_t b = b0 + col;
        if (b >= B) continue;
        const float m = mrun[col];
#pragma unroll
        for (int a = 0; a < RM; ++a) {
            const int64_t o = (crow0 + ty + 16 * a) * B + b;
            const float e = emars[o];
            float v = (m == kNegInf || e == kNegInf) ? 0.0f : acc[a][c] * __expf(e + m);
            if (accumulate) v += eflows[o];
            eflows[o] = v;
        }
    }
}

// Element flows for child blocks smaller than 16: one thread per (child node, batch column).
__global__ __launch_bounds__(kThreads) void sum_bwd_ele_rowwise(float* __restrict__ eflows, const float* __restrict__ nflows,
                                                                const float* __restrict__ nmars, const float* __restrict__ emars,
                                                                const float* __restrict__ params, const int* __restrict__ chids,
                                                                const int* __restrict__ parids, const int* __restrict__ parpids,
                                                                const int* __restrict__ npar, int KP, int cbs, int bs,
                                                                int64_t B, bool modified, bool accumulate) {
    const int64_t b = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int j = blockIdx.x / cbs, ci = blockIdx.x % cbs;
    const int kflat = npar[j] * bs;
    const int* parrow = parids + (int64_t)j * KP;
    const int* pprow = parpids + (int64_t)j * KP;
    float m = kNegInf;
    for (int kk = 0; kk < kflat; ++kk) {
        const int kb = kk / bs, pi = kk - kb * bs;
        m = fmaxf(m, log_flow_div_mar(nflows, nmars, ((int64_t)parrow[kb] + pi) * B + b, modified));
    }
    const int64_t o = ((int64_t)chids[j] + ci) * B + b;
    const float e = emars[o];
    float v = 0.0f;
    if (m != kNegInf && e != kNegInf) {
        float s = 0.0f;
        for (int kk = 0; kk < kflat; ++kk) {
            const int kb = kk / bs, pi = kk - kb * bs;
            const float l = log_flow_div_mar(nflows, nmars, ((int64_t)parrow[kb] + pi) * B + b, modified);
            s = fmaf(params[(int64_t)pprow[kb] + (int64_t)ci * bs + pi], __expf(l - m), s);
        }
        v = s * __expf(e + m);
    }
    if (accumulate) v += eflows[o];
    eflows[o] = v;
}

// ---------------------------------------------------------------------------------------------
// Parameter flows: pf[pfid_k + i] += theta[pid_k + i] * sum_b exp(lfdm[nid + i, b] + e[cid_k, b]).
// Workgroup = (node block j, TM parent rows) x (TN children); loops over the whole batch in
// chunks of TB, shifting every batch column by the chunk's max over the TM rows.
template <int TM, int TN, typename PF>
__global__ __launch_bounds__(kThreads) void sum_bwd_par_tiled(PF* __restrict__ pflows, const float* __restrict__ nflows,
                                                              const float* __restrict__ nmars, const float* __restrict__ emars,
                                                              const float* __restrict__ params, const int* __restrict__ nids,
                                                              const int* __restrict__ cids, const int* __restrict__ pids,
                                                              const int* __restrict__ pfids, const int* __restrict__ nch,
                                                              int K, int bs, int64_t B, bool modified, int atomic_out,
                                                              float sign) {
    constexpr int TB = 32;
    constexpr int RM = TM / 16, RN = TN / 16;
    __shared__ float Xs[TB][TM + 4];
    __shared__ float Ys[TB][TN + 4];
    __shared__ float shift[TB];

    const int mtiles = bs / TM;
    const int j = blockIdx.x / mtiles;
    const int m0 = (blockIdx.x % mtiles) * TM;
    const int k0 = blockIdx.y * TN;
    const int n = nch[j];
    if (k0 >= n) return;
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int64_t nrow0 = (int64_t)nids[j] + m0;
    const int* crow = cids + (int64_t)j * K;

    // loaders: X tile = TM rows x TB columns (lanes along b); Y tile = TN rows x TB columns
    const int lb = tid % TB, lr = tid / TB;   // 8 rows per pass
    constexpr int PASS = kThreads / TB;

    float acc[RM][RN];
#pragma unroll
    for (int a = 0; a < RM; ++a)
#pragma unroll
        for (int c = 0; c < RN; ++c) acc[a][c] = 0.0f;

    for (int64_t bb = 0; bb < B; bb += TB) {
        const int64_t b = bb + lb;
        float xv[TM / PASS];
        float lmax = kNegInf;
#pragma unroll
        for (int r = 0; r < TM / PASS; ++r) {
            const int row = lr + r * PASS;
            float v = kNegInf;
            if (b < B) v = log_flow_div_mar(nflows, nmars, (nrow0 + row) * B + b, modified);
            xv[r] = v;
            lmax = fmaxf(lmax, v);
        }
        // column max over the TM rows: reduce across the PASS row groups that share lane lb
        for (int o = TB; o < 64; o <<= 1) lmax = fmaxf(lmax, __shfl_xor(lmax, o, 64));
        __shared__ float wred[kThreads / 64][TB];
        if ((tid & 63) < TB) wred[tid >> 6][lb] = lmax;
        __syncthreads();
        if (tid < TB) {
            float t = kNegInf;
            for (int q = 0; q < kThreads / 64; ++q) t = fmaxf(t, wred[q][tid]);
            shift[tid] = t;
        }
        __syncthreads();
        const float sh = shift[lb];
#pragma unroll
        for (int r = 0; r < TM / PASS; ++r) Xs[lb][lr + r * PASS] = (sh == kNegInf) ? 0.0f : __expf(xv[r] - sh);
#pragma unroll
        for (int r = 0; r < TN / PASS; ++r) {
            const int kk = lr + r * PASS;
            const int k = k0 + kk;
            float v = 0.0f;
            if (k < n && b < B && sh != kNegInf) {
                const float e = emars[(int64_t)crow[k] * B + b];
                v = (e == kNegInf) ? 0.0f : __expf(e + sh);
            }
            Ys[lb][kk] = v;
        }
        __syncthreads();
#pragma unroll 8
        for (int t = 0; t < TB; ++t) {
            float av[RM], bv[RN];
#pragma unroll
            for (int a = 0; a < RM; ++a) av[a] = Xs[t][ty + 16 * a];
#pragma unroll
            for (int c = 0; c < RN; ++c) bv[c] = Ys[t][tx + 16 * c];
#pragma unroll
            for (int a = 0; a < RM; ++a)
#pragma unroll
                for (int c = 0; c < RN; ++c) acc[a][c] = fmaf(av[a], bv[c], acc[a][c]);
        }
        __syncthreads();
    }
    const int* prow = pids + (int64_t)j * K;
    const int* pfrow = pfids + (int64_t)j * K;
#pragma unroll
    for (int c = 0; c < RN; ++c) {
        const int k = k0 + tx + 16 * c;
        if (k >= n) continue;
        const int64_t p0 = (int64_t)prow[k] + m0, f0 = (int64_t)pfrow[k] + m0;
#pragma unroll
        for (int a = 0; a < RM; ++a) {
            const int i = ty + 16 * a;
            const float v = sign * params[p0 + i] * acc[a][c];
            pf_add<PF>(&pflows[f0 + i], (PF)v, atomic_out);
        }
    }
}

// Parameter flows for bs < 16: one wavefront per (node, child), lanes along the batch (loop
// unrolled so that several batch chunks are in flight), then a wavefront reduction.
template <typename PF>
__global__ __launch_bounds__(kThreads) void sum_bwd_par_rowwise(PF* __restrict__ pflows, const float* __restrict__ nflows,
                                                                const float* __restrict__ nmars, const float* __restrict__ emars,
                                                                const float* __restrict__ params, const int* __restrict__ nids,
                                                                const int* __restrict__ cids, const int* __restrict__ pids,
                                                                const int* __restrict__ pfids, const int* __restrict__ nch,
                                                                int K, int bs, int64_t B, bool modified, int atomic_out,
                                                                float sign) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = blockIdx.y * kSmallWaves + w;
    const int j = blockIdx.x / bs, i = blockIdx.x % bs;
    if (k >= nch[j]) return;
    const int64_t nrow = (int64_t)nids[j] + i;
    const int64_t erow = cids[(int64_t)j * K + k];
    float s = 0.0f;
#pragma unroll 4
    for (int64_t b = lane; b < B; b += 64) {
        const float l = log_flow_div_mar(nflows, nmars, nrow * B + b, modified);
        const float e = emars[erow * B + b];
        if (l != kNegInf && e != kNegInf) s += __expf(l + e);
    }
    s = warp_sum(s);
    if (lane == 0) {
        const float v = sign * params[(int64_t)pids[(int64_t)j * K + k] + i] * s;
        pf_add<PF>(&pflows[(int64_t)pfids[(int64_t)j * K + k] + i], (PF)v, atomic_out);
    }
}


// ---------------------------------------------------------------------------------------------
// Other propagation algorithms (sum_layer.py:454-800 forward, 1352-1392 modify pass, 1569-1730
// element flows, 2049-2169 parameter flows), on the vector ALUs: alg 1 = MPE, 2 = GeneralLL with
// exponent alpha; alg 0 (LL) is GeneralLL with alpha = 1.  lt = log theta + e is formed the
// same way (same operations in the same order) in the forward and in both backward kernels, so
// MPE's match test |lt - m| < 1e-6 finds the forward's own maximum.

// node_mars[nid + i][b]: MPE max_k lt_k; GeneralLL log(sum_k exp(alpha lt_k)) / alpha.  One thread
// per (node row, column).
__global__ __launch_bounds__(kThreads) void sum_fwd_alg_kernel(float* __restrict__ node_mars,
                                                               const float* __restrict__ emars,
                                                               const float* __restrict__ params,
                                                               const int* __restrict__ nids,
                                                               const int* __restrict__ cids,
                                                               const int* __restrict__ pids,
                                                               const int* __restrict__ nch, int K, int bs, int64_t B,
                                                               int alg, float alpha) {
    const int64_t b = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int j = blockIdx.x / bs, i = blockIdx.x % bs;
    const int n = nch[j];
    const int* crow = cids + (int64_t)j * K;
    const int* prow = pids + (int64_t)j * K;
    float m = kNegInf, s = 0.0f;
    for (int k = 0; k < n; ++k) {
        const float lt = __logf(params[(int64_t)prow[k] + i]) + emars[(int64_t)crow[k] * B + b];
        if (alg == 1) {
            m = fmaxf(m, lt);
        } else {
            const float x = alpha * lt;
            if (x == kNegInf) continue;
            if (x > m) {
                s = s * __expf(m - x) + 1.0f;
                m = x;
            } else {
                s += __expf(x - m);
            }
        }
    }
    float v = m;
    if (alg != 1) v = (m == kNegInf) ? kNegInf : (m + __logf(s)) / alpha;
    node_mars[((int64_t)nids[j] + i) * B + b] = v;
}

// element_flows[chid + ci][b] = sum over parents p of: MPE nf[p][b] where |lt - nm[p][b]| < 1e-6;
// GeneralLL nf[p][b] exp(alpha (lt - nm[p][b])), the child's share of the parent's power sum.
// Flows are read as stored (no in-place transform for these algorithms).  One thread per (child
// row, column); accumulate adds to the element flows instead of storing.
__global__ __launch_bounds__(kThreads) void sum_bwd_ele_alg_kernel(float* __restrict__ eflows,
                                                                   const float* __restrict__ nflows,
                                                                   const float* __restrict__ nmars,
                                                                   const float* __restrict__ emars,
                                                                   const float* __restrict__ params,
                                                                   const int* __restrict__ chids,
                                                                   const int* __restrict__ parids,
                                                                   const int* __restrict__ parpids,
                                                                   const int* __restrict__ npar, int KP, int cbs, int bs,
                                                                   int64_t B, int alg, float alpha, bool accumulate) {
    const int64_t b = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int j = blockIdx.x / cbs, ci = blockIdx.x % cbs;
    const int kflat = npar[j] * bs;
    const int* parrow = parids + (int64_t)j * KP;
    const int* pprow = parpids + (int64_t)j * KP;
    const int64_t o = ((int64_t)chids[j] + ci) * B + b;
    const float e = emars[o];
    float acc = 0.0f;
    for (int kk = 0; kk < kflat; ++kk) {
        const int kb = kk / bs, pi = kk - kb * bs;
        const int64_t po = ((int64_t)parrow[kb] + pi) * B + b;
        const float nf = nflows[po], nm = nmars[po];
        const float lt = __logf(params[(int64_t)pprow[kb] + (int64_t)ci * bs + pi]) + e;
        if (alg == 1) {
            if (fabsf(lt - nm) < 1e-6f) acc += nf;
        } else if (nm != kNegInf && lt != kNegInf && nf != 0.0f) {
            acc += nf * __expf(alpha * (lt - nm));
        }
    }
    if (accumulate) acc += eflows[o];
    eflows[o] = acc;
}

// pf[pfid_k + i] += MPE: sum_b nf[b] where |lt_k - nm[b]| < 1e-6 (not scaled by theta);
// GeneralLL: theta_k sum_b nf[b] exp(e_k[b] - nm[b]) (the LL form, as the reference computes it).
// One wavefront per (node row, child), lanes along the batch; slots added atomically.
template <typename PF>
__global__ __launch_bounds__(kThreads) void sum_bwd_par_alg_kernel(PF* __restrict__ pflows,
                                                                   const float* __restrict__ nflows,
                                                                   const float* __restrict__ nmars,
                                                                   const float* __restrict__ emars,
                                                                   const float* __restrict__ params,
                                                                   const int* __restrict__ nids,
                                                                   const int* __restrict__ cids,
                                                                   const int* __restrict__ pids,
                                                                   const int* __restrict__ pfids,
                                                                   const int* __restrict__ nch, int K, int bs,
                                                                   int64_t B, int alg, float sign) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = blockIdx.y * (kThreads / 64) + w;
    const int j = blockIdx.x / bs, i = blockIdx.x % bs;
    if (k >= nch[j]) return;
    const int64_t nrow = (int64_t)nids[j] + i;
    const int64_t erow = cids[(int64_t)j * K + k];
    const float th = params[(int64_t)pids[(int64_t)j * K + k] + i];
    const float lth = __logf(th);
    float s = 0.0f;
    for (int64_t b = lane; b < B; b += 64) {
        const float nf = nflows[nrow * B + b], nm = nmars[nrow * B + b];
        const float e = emars[erow * B + b];
        if (alg == 1) {
            if (fabsf(lth + e - nm) < 1e-6f) s += nf;
        } else if (nm != kNegInf && e != kNegInf && nf != 0.0f) {
            s += nf * __expf(e - nm);
        }
    }
    s = warp_sum(s);
    if (lane == 0) {
        const float v = sign * (alg == 1 ? s : th * s);
        if (v != 0.0f) atomicAdd(&pflows[(int64_t)pfids[(int64_t)j * K + k] + i], (PF)v);
    }
}
}  // namespace

// precision: 0 = FP32 operands on MFMA, 1 = BF16 MFMA, 2 = FP16 MFMA (pjamd_sum_mfma.hip, block sizes
// >= 32); 3 = FP32 on the vector ALUs (this file).  Smaller blocks always use this file's kernels.
int pjamd_sum_forward_mfma(float* node_mars, const float* element_mars, const float* params, const int* nids,
                           const int* cids, const int* pids, const int* nch, int n_blocks, int K, int bs, int64_t B,
                           int child_run, int precision, hipStream_t stream);
int pjamd_sum_backward_ele_mfma(float* eflows, const float* nflows, const float* nmars, const float* emars,
                                const float* params, const int* chids, const int* parids, const int* parpids,
                                const int* npar, int n, int KP, int cbs, int bs, int64_t B, bool mod, bool accum,
                                bool aligned, int precision, hipStream_t stream);
int pjamd_sum_backward_par_mfma(void* pflows, bool pf64, const float* nflows, const float* nmars, const float* emars,
                                const float* params, const int* nids, const int* cids, const int* pids, const int* pfids,
                                const int* nch, int n_blocks, int K, int bs, int64_t B, bool mod, int pf_mode, float sign,
                                int precision, hipStream_t stream);

PJAMD_EXPORT int pjamd_sum_forward(float* node_mars, const float* element_mars, const float* params, const int* nids,
                                   const int* cids, const int* pids, const int* nch, int n_blocks, int K, int bs,
                                   int64_t B, int child_run, int precision, hipStream_t stream) {
    PJAMD_REQUIRE(n_blocks >= 0 && K >= 1 && bs >= 1 && B > 0 && precision >= 0 && precision <= 3 && child_run >= 0);
    if (n_blocks == 0) return PJAMD_OK;
    if (precision != 3 && bs >= 32)
        return pjamd_sum_forward_mfma(node_mars, element_mars, params, nids, cids, pids, nch, n_blocks, K, bs, B,
                                      child_run, precision, stream);
    const unsigned gy = cdiv(B, kTN);
    if (bs >= 64) {
        sum_fwd_tiled<64><<<dim3((unsigned)((int64_t)n_blocks * (bs / 64)), gy), kThreads, 0, stream>>>(
            node_mars, element_mars, params, nids, cids, pids, nch, K, bs, B);
    } else if (bs == 32) {
        sum_fwd_tiled<32><<<dim3((unsigned)n_blocks, gy), kThreads, 0, stream>>>(node_mars, element_mars, params, nids,
                                                                                cids, pids, nch, K, bs, B);
    } else if (bs == 16) {
        sum_fwd_tiled<16><<<dim3((unsigned)n_blocks, gy), kThreads, 0, stream>>>(node_mars, element_mars, params, nids,
                                                                                cids, pids, nch, K, bs, B);
    } else {
        // wide rows (a root over hundreds of children): 16 wavefronts share a node's children
        const dim3 grid((unsigned)((int64_t)n_blocks * bs), cdiv(B, 64));
        if (K >= 8 * kRowChunk)
            sum_fwd_rowwise<16><<<grid, 1024, 0, stream>>>(node_mars, element_mars, params, nids, cids, pids, nch, K,
                                                           bs, B);
        else
            sum_fwd_rowwise<4><<<grid, 256, 0, stream>>>(node_mars, element_mars, params, nids, cids, pids, nch, K, bs,
                                                         B);
    }
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_sum_modify_flows(float* node_flows, const float* node_mars, const int* nids, int n_blocks, int bs,
                                        int64_t B, hipStream_t stream) {
    PJAMD_REQUIRE(n_blocks >= 0 && bs >= 1 && B > 0);
    if (n_blocks == 0) return PJAMD_OK;
    modify_flows_kernel<<<dim3((unsigned)((int64_t)n_blocks * bs), cdiv(B, kThreads)), kThreads, 0, stream>>>(
        node_flows, node_mars, nids, bs, B);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

// flags: bit 0 = node flows hold log(flow) - mar; bit 1 = add to existing element flows.
PJAMD_EXPORT int pjamd_sum_backward_ele(float* element_flows, const float* node_flows, const float* node_mars,
                                        const float* element_mars, const float* params, const int* chids,
                                        const int* parids, const int* parpids, const int* npar, int n_cblocks, int KP,
                                        int cbs, int bs, int64_t B, int flags, int precision, hipStream_t stream) {
    PJAMD_REQUIRE(n_cblocks >= 0 && KP >= 1 && cbs >= 1 && bs >= 1 && B > 0 && precision >= 0 && precision <= 3);
    if (n_cblocks == 0) return PJAMD_OK;
    const bool mod = flags & 1, accum = (flags & 2) != 0, aligned = (flags & 4) != 0;
    if (precision != 3 && cbs >= 32)
        return pjamd_sum_backward_ele_mfma(element_flows, node_flows, node_mars, element_mars, params, chids, parids,
                                           parpids, npar, n_cblocks, KP, cbs, bs, B, mod, accum, aligned, precision,
                                           stream);
    const unsigned gy = cdiv(B, kTN);
    if (cbs >= 64) {
        sum_bwd_ele_tiled<64><<<dim3((unsigned)((int64_t)n_cblocks * (cbs / 64)), gy), kThreads, 0, stream>>>(
            element_flows, node_flows, node_mars, element_mars, params, chids, parids, parpids, npar, KP, cbs, bs, B,
            mod, accum);
    } else if (cbs == 32) {
        sum_bwd_ele_tiled<32><<<dim3((unsigned)n_cblocks, gy), kThreads, 0, stream>>>(
            element_flows, node_flows, node_mars, element_mars, params, chids, parids, parpids, npar, KP, cbs, bs, B,
            mod, accum);
    } else if (cbs == 16) {
        sum_bwd_ele_tiled<16><<<dim3((unsigned)n_cblocks, gy), kThreads, 0, stream>>>(
            element_flows, node_flows, node_mars, element_mars, params, chids, parids, parpids, npar, KP, cbs, bs, B,
            mod, accum);
    } else {
        sum_bwd_ele_rowwise<<<dim3((unsigned)((int64_t)n_cblocks * cbs), cdiv(B, kThreads)), kThreads, 0, stream>>>(
            element_flows, node_flows, node_mars, element_mars, params, chids, parids, parpids, npar, KP, cbs, bs, B,
            mod, accum);
    }
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

template <typename PF>
static int sum_bwd_par_impl(PF* pflows, const float* node_flows, const float* node_mars, const float* element_mars,
                            const float* params, const int* nids, const int* cids, const int* pids, const int* pfids,
                            const int* nch, int n_blocks, int K, int bs, int64_t B, bool mod, int atomic_out,
                            float sign, hipStream_t stream) {
    if (bs >= 64) {
        sum_bwd_par_tiled<64, 64, PF><<<dim3((unsigned)((int64_t)n_blocks * (bs / 64)), cdiv(K, 64)), kThreads, 0, stream>>>(
            pflows, node_flows, node_mars, element_mars, params, nids, cids, pids, pfids, nch, K, bs, B, mod,
            atomic_out, sign);
    } else if (bs == 32) {
        sum_bwd_par_tiled<32, 64, PF><<<dim3((unsigned)n_blocks, cdiv(K, 64)), kThreads, 0, stream>>>(
            pflows, node_flows, node_mars, element_mars, params, nids, cids, pids, pfids, nch, K, bs, B, mod,
            atomic_out, sign);
    } else if (bs == 16) {
        sum_bwd_par_tiled<16, 64, PF><<<dim3((unsigned)n_blocks, cdiv(K, 64)), kThreads, 0, stream>>>(
            pflows, node_flows, node_mars, element_mars, params, nids, cids, pids, pfids, nch, K, bs, B, mod,
            atomic_out, sign);
    } else {
        sum_bwd_par_rowwise<PF><<<dim3((unsigned)((int64_t)n_blocks * bs), cdiv(K, kSmallWaves)), kThreads, 0, stream>>>(
            pflows, node_flows, node_mars, element_mars, params, nids, cids, pids, pfids, nch, K, bs, B, mod,
            atomic_out, sign);
    }
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

// flags: bit 0 = node flows hold log(flow) - mar; bit 1 = flow slots shared inside the launch
// (atomic adds); bit 2 = subtract instead of add; bit 3 = parameter flows are FP64; bit 4 = the
// accumulator is zero: store instead of read-modify-write (not with bit 1).
PJAMD_EXPORT int pjamd_sum_backward_par(void* param_flows, const float* node_flows, const float* node_mars,
                                        const float* element_mars, const float* params, const int* nids, const int* cids,
                                        const int* pids, const int* pfids, const int* nch, int n_blocks, int K, int bs,
                                        int64_t B, int flags, int precision, hipStream_t stream) {
    PJAMD_REQUIRE(n_blocks >= 0 && K >= 1 && bs >= 1 && B > 0 && precision >= 0 && precision <= 3);
    if (n_blocks == 0) return PJAMD_OK;
    PJAMD_REQUIRE(!((flags & 2) && (flags & 16)));
    const bool mod = flags & 1;
    const int atomic_out = (flags & 2) ? kPfAtomic : (flags & 16) ? kPfOverwrite : kPfAdd;
    const float sign = (flags & 4) ? -1.0f : 1.0f;
    if (precision != 3 && bs >= 32)
        return pjamd_sum_backward_par_mfma(param_flows, (flags & 8) != 0, node_flows, node_mars, element_mars, params,
                                           nids, cids, pids, pfids, nch, n_blocks, K, bs, B, mod, atomic_out, sign,
                                           precision, stream);
    if (flags & 8)
        return sum_bwd_par_impl<double>((double*)param_flows, node_flows, node_mars, element_mars, params, nids, cids,
                                        pids, pfids, nch, n_blocks, K, bs, B, mod, atomic_out, sign, stream);
    return sum_bwd_par_impl<float>((float*)param_flows, node_flows, node_mars, element_mars, params, nids, cids, pids,
                                   pfids, nch, n_blocks, K, bs, B, mod, atomic_out, sign, stream);
}

// alg: 0 = LL (as GeneralLL with alpha = 1), 1 = MPE, 2 = GeneralLL.
PJAMD_EXPORT int pjamd_sum_forward_alg(float* node_mars, const float* element_mars, const float* params,
                                       const int* nids, const int* cids, const int* pids, const int* nch,
                                       int n_blocks, int K, int bs, int64_t B, int alg, float alpha,
                                       hipStream_t stream) {
    PJAMD_REQUIRE(n_blocks >= 0 && K >= 1 && bs >= 1 && B > 0 && alg >= 0 && alg <= 2 &&
                  (int64_t)n_blocks * bs < ((int64_t)1 << 31));
    PJAMD_REQUIRE(alg != 2 || alpha > 0.0f);
    if (n_blocks == 0) return PJAMD_OK;
    PJAMD_REQUIRE(cdiv(B, kThreads) < 65536u);
    dim3 grid((unsigned)(n_blocks * bs), cdiv(B, kThreads));
    sum_fwd_alg_kernel<<<grid, kThreads, 0, stream>>>(node_mars, element_mars, params, nids, cids, pids, nch, K, bs, B,
                                                      alg, alg == 0 ? 1.0f : alpha);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_sum_backward_ele_alg(float* element_flows, const float* node_flows, const float* node_mars,
                                            const float* element_mars, const float* params, const int* chids,
                                            const int* parids, const int* parpids, const int* n_pars, int n_cblocks,
                                            int KP, int cbs, int bs, int64_t B, int alg, float alpha, int accumulate,
                                            hipStream_t stream) {
    PJAMD_REQUIRE(n_cblocks >= 0 && KP >= 1 && cbs >= 1 && bs >= 1 && B > 0 && alg >= 0 && alg <= 2 &&
                  (int64_t)n_cblocks * cbs < ((int64_t)1 << 31));
    PJAMD_REQUIRE(alg != 2 || alpha > 0.0f);
    if (n_cblocks == 0) return PJAMD_OK;
    PJAMD_REQUIRE(cdiv(B, kThreads) < 65536u);
    dim3 grid((unsigned)(n_cblocks * cbs), cdiv(B, kThreads));
    sum_bwd_ele_alg_kernel<<<grid, kThreads, 0, stream>>>(element_flows, node_flows, node_mars, element_mars, params,
                                                          chids, parids, parpids, n_pars, KP, cbs, bs, B, alg,
                                                          alg == 0 ? 1.0f : alpha, accumulate != 0);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_sum_backward_par_alg(void* param_flows, int pf64, const float* node_flows,
                                            const float* node_mars, const float* element_mars, const float* params,
                                            const int* nids, const int* cids, const int* pids, const int* pfids,
                                            const int* nch, int n_blocks, int K, int bs, int64_t B, int alg,
                                            hipStream_t stream) {
    const float sign = (alg & PJAMD_ALG_NEGATE) ? -1.0f : 1.0f;
    alg &= ~PJAMD_ALG_NEGATE;
    PJAMD_REQUIRE(n_blocks >= 0 && K >= 1 && bs >= 1 && B > 0 && alg >= 0 && alg <= 2 &&
                  (int64_t)n_blocks * bs < ((int64_t)1 << 31));
    PJAMD_REQUIRE(cdiv(K, kThreads / 64) < 65536u);
    if (n_blocks == 0) return PJAMD_OK;
    dim3 grid((unsigned)(n_blocks * bs), cdiv(K, kThreads / 64));
    if (pf64)
        sum_bwd_par_alg_kernel<double><<<grid, kThreads, 0, stream>>>((double*)param_flows, node_flows, node_mars,
                                                                      element_mars, params, nids, cids, pids, pfids,
                                                                      nch, K, bs, B, alg, sign);
    else
        sum_bwd_par_alg_kernel<float><<<grid, kThreads, 0, stream>>>((float*)param_flows, node_flows, node_mars,
                                                                     element_mars, params, nids, cids, pids, pfids,
                                                                     nch, K, bs, B, alg, sign);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}
