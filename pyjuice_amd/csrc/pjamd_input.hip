// Input layers: Categorical (kind 0) and Bernoulli (kind 1) forward, flow backward,
// mini-batch EM and sampling.  Data is [num_vars, B] int32; node rows are [num_nodes, B].
#include "pjamd_common.h"

#include <atomic>

using namespace pjamd;

namespace {

constexpr int kLdsFloats = 4096; // largest category count with LDS flow histograms

// node_mars[(off + i) * B + b] = log p_i(x[v_i, b]); categories past K have p = 0, negative ones
// (missing entries) are marginalized: log 1 = 0.
// kFwdRowThreads threads per group of NT consecutive node rows, 4 consecutive batch columns each
// per pass (16-byte loads of the categories, 16-byte stores).  A thread issues the category loads
// of all NT nodes, then all 4 NT parameter gathers (they hit each node's row of kmax floats in
// L1 / L2), so a workgroup waits for two memory round trips per NT rows instead of per row.
// Measured at 200704 nodes x 512 columns (ms): 128 threads x 4 nodes 0.178, x 2 0.179, x 8 0.199;
// 64 x 2 0.165, 64 x 4 0.168, 64 x 8 0.185; 32 x 2 0.168, 32 x 4 0.171.
constexpr int kFwdRowThreads = 64;
constexpr int kFwdNodes = 2;
constexpr int kRowThreads = 128;   // the LDS-histogram backward's threads per node

// log p of one data entry.  A negative category marks a marginalized (missing) entry, whose
// probability summed over all values is 1; a category past K has probability 0.  th is the
// gathered parameter of category x (any finite value when x is out of range).
__device__ __forceinline__ float cat_logp(int x, int K, float th) {
    return x < 0 ? 0.0f : (x < K ? __logf(th) : kNegInf);
}
__device__ __forceinline__ float bern_logp(int x, float p1) {
    return x < 0 ? 0.0f : __logf(x > 0 ? p1 : 1.0f - p1);
}

template <int KIND, int NT>
__global__ __launch_bounds__(256) void input_fwd_kernel(const float* __restrict__ params, float* __restrict__ node_mars,
                                                        const int* __restrict__ data, const int* __restrict__ vids,
                                                        const int64_t* __restrict__ s_pids,
                                                        const int* __restrict__ ncats, int64_t num_nodes,
                                                        int64_t node_offset, int64_t B) {
    const int64_t i0 = ((int64_t)blockIdx.x * (256 / kFwdRowThreads) + threadIdx.x / kFwdRowThreads) * NT;
    if (i0 >= num_nodes) return;
    const int nn = (num_nodes - i0) < NT ? (int)(num_nodes - i0) : NT;
    const int sub = threadIdx.x % kFwdRowThreads;
    const int* xrow[NT];
    const float* th[NT];
    int K[NT];
    float p1[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int64_t i = i0 + (n < nn ? n : 0);    // rows past the layer repeat row i0, never stored
        xrow[n] = data + (int64_t)vids[i] * B;
        th[n] = params + s_pids[i];
        K[n] = (KIND == 0) ? ncats[i] : 2;
        p1[n] = (KIND == 1) ? th[n][0] : 0.0f;
    }
    for (int64_t b = (int64_t)sub * 4; b < B; b += kFwdRowThreads * 4) {
        if (b + 4 <= B) {
            int4 x[NT];
#pragma unroll
            for (int n = 0; n < NT; ++n) x[n] = *reinterpret_cast<const int4*>(xrow[n] + b);
            float4 o[NT];
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                if (KIND == 0) {
                    const int k = K[n];
                    const int c0 = (x[n].x >= 0 && x[n].x < k) ? x[n].x : 0;
                    const int c1 = (x[n].y >= 0 && x[n].y < k) ? x[n].y : 0;
                    const int c2 = (x[n].z >= 0 && x[n].z < k) ? x[n].z : 0;
                    const int c3 = (x[n].w >= 0 && x[n].w < k) ? x[n].w : 0;
                    o[n] = make_float4(th[n][c0], th[n][c1], th[n][c2], th[n][c3]);
                }
            }
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                float4 r;
                if (KIND == 0) {
                    const int k = K[n];
                    r.x = cat_logp(x[n].x, k, o[n].x);
                    r.y = cat_logp(x[n].y, k, o[n].y);
                    r.z = cat_logp(x[n].z, k, o[n].z);
                    r.w = cat_logp(x[n].w, k, o[n].w);
                } else {
                    r.x = bern_logp(x[n].x, p1[n]);
                    r.y = bern_logp(x[n].y, p1[n]);
                    r.z = bern_logp(x[n].z, p1[n]);
                    r.w = bern_logp(x[n].w, p1[n]);
                }
                if (n < nn) *reinterpret_cast<float4*>(node_mars + (node_offset + i0 + n) * B + b) = r;
            }
        } else {
            for (int n = 0; n < nn; ++n) {
                float* out = node_mars + (node_offset + i0 + n) * B;
                for (int64_t c = b; c < B; ++c) {
                    const int xx = xrow[n][c];
                    if (KIND == 0) out[c] = cat_logp(xx, K[n], (xx >= 0 && xx < K[n]) ? th[n][xx] : 1.0f);
                    else out[c] = bern_logp(xx, p1[n]);
                }
            }
        }
    }
}

// Categorical forward with each node's log-parameters staged in LDS (every category count <=
// kFwdLdsK, B % 4 == 0): a wavefront owns NT consecutive nodes, loads their parameter rows with
// coalesced loads issued beside its first category loads (no load waits for another), takes
// each parameter's log once instead of once per sample, and reads log p from its own LDS rows.
// There is no workgroup barrier: a wavefront reads only the rows it wrote, after its LDS writes
// have completed (lgkmcnt 0).
constexpr int kFwdLdsK = 256;

template <int NT>
__global__ __launch_bounds__(256) void input_fwd_lds_kernel(const float* __restrict__ params,
                                                            float* __restrict__ node_mars,
                                                            const int* __restrict__ data, const int* __restrict__ vids,
                                                            const int64_t* __restrict__ s_pids,
                                                            const int* __restrict__ ncats, int64_t num_nodes,
                                                            int64_t node_offset, int64_t B) {
    __shared__ float lt[4][NT][kFwdLdsK];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t i0 = ((int64_t)blockIdx.x * 4 + w) * NT;
    if (i0 >= num_nodes) return;   // the whole wavefront
    const int nn = (num_nodes - i0) < NT ? (int)(num_nodes - i0) : NT;
    const int* xrow[NT];
    int K[NT];
    float pv[NT][kFwdLdsK / 64];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int64_t i = i0 + (n < nn ? n : 0);    // rows past the layer repeat row i0, never stored
        xrow[n] = data + (int64_t)vids[i] * B;
        K[n] = ncats[i];
        const float* th = params + s_pids[i];
#pragma unroll
        for (int u = 0; u < kFwdLdsK / 64; ++u) {
            const int k = lane + 64 * u;
            pv[n][u] = th[k < K[n] ? k : 0];
        }
    }
    int64_t b = 4 * (int64_t)lane;
    int4 xc[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) xc[n] = b < B ? *reinterpret_cast<const int4*>(xrow[n] + b) : make_int4(0, 0, 0, 0);
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
        for (int u = 0; u < kFwdLdsK / 64; ++u) lt[w][n][lane + 64 * u] = __logf(pv[n][u]);
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    for (; b < B; b += 256) {
        const int64_t bn = b + 256;
        int4 xn[NT];
#pragma unroll
        for (int n = 0; n < NT; ++n)
            xn[n] = bn < B ? *reinterpret_cast<const int4*>(xrow[n] + bn) : make_int4(0, 0, 0, 0);
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const int k = K[n] < kFwdLdsK ? K[n] : kFwdLdsK;   // the caller promises K <= kmax <= kFwdLdsK
            const float* row = lt[w][n];
            auto lp = [&](int x) { return x < 0 ? 0.0f : (x < k ? row[x] : kNegInf); };
            const float4 r = make_float4(lp(xc[n].x), lp(xc[n].y), lp(xc[n].z), lp(xc[n].w));
            if (n < nn) *reinterpret_cast<float4*>(node_mars + (node_offset + i0 + n) * B + b) = r;
        }
#pragma unroll
        for (int n = 0; n < NT; ++n) xc[n] = xn[n];
    }
}

// Forward, generic form (any B and alignment): one thread per (node, batch column).
__global__ __launch_bounds__(256) void cat_fwd_global_kernel(const float* __restrict__ params,
                                                             float* __restrict__ node_mars,
                                                             const int* __restrict__ data, const int* __restrict__ vids,
                                                             const int64_t* __restrict__ s_pids,
                                                             const int* __restrict__ ncats, int64_t node_offset,
                                                             int64_t B, int kind) {
    const int64_t b = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int64_t i = blockIdx.x;
    const int x = data[(int64_t)vids[i] * B + b];
    float v;
    if (kind == 0) {
        v = cat_logp(x, ncats[i], (x >= 0 && x < ncats[i]) ? params[s_pids[i] + x] : 1.0f);
    } else {
        v = bern_logp(x, params[s_pids[i]]);
    }
    node_mars[(node_offset + i) * B + b] = v;
}

// Categorical flow backward: a workgroup owns two nodes (kRowThreads threads each, 4 consecutive
// batch columns per thread with 16-byte loads) and builds their flow histograms in LDS; then
// each node's kmax slots are added to the parameter flows, all slot reads issued before any
// write (the slots of a node are distinct), or atomically when slots are shared between nodes.
template <typename PF>
__global__ __launch_bounds__(256) void cat_bwd_lds_kernel(PF* __restrict__ pflows, const float* __restrict__ node_flows,
                                                          const int* __restrict__ data, const int* __restrict__ vids,
                                                          const int64_t* __restrict__ s_pfids,
                                                          const int* __restrict__ ncats, int64_t num_nodes,
                                                          int64_t node_offset, int64_t B, int kmax, int pf_mode) {
    extern __shared__ float hist[];
    constexpr int NR = 256 / kRowThreads;
    const int r = threadIdx.x / kRowThreads, sub = threadIdx.x % kRowThreads;
    const int64_t i = (int64_t)blockIdx.x * NR + r;
    const bool on = i < num_nodes;
    float* h = hist + r * kmax;
    for (int t = sub; t < kmax; t += kRowThreads) h[t] = 0.0f;
    __syncthreads();
    const int K = on ? ncats[i] : 0;
    if (on) {
        const int* xrow = data + (int64_t)vids[i] * B;
        const float* frow = node_flows + (node_offset + i) * B;
        for (int64_t b = (int64_t)sub * 4; b < B; b += kRowThreads * 4) {
            const int4 x = *reinterpret_cast<const int4*>(xrow + b);
            const float4 f = *reinterpret_cast<const float4*>(frow + b);
            if (x.x >= 0 && x.x < K && f.x != 0.0f) atomicAdd(&h[x.x], f.x);
            if (x.y >= 0 && x.y < K && f.y != 0.0f) atomicAdd(&h[x.y], f.y);
            if (x.z >= 0 && x.z < K && f.z != 0.0f) atomicAdd(&h[x.z], f.z);
            if (x.w >= 0 && x.w < K && f.w != 0.0f) atomicAdd(&h[x.w], f.w);
        }
    }
    __syncthreads();
    if (!on) return;
    constexpr int kPer = 8;   // slots per thread per pass (loads first)
    const int64_t f0 = s_pfids[i];
    for (int k0 = 0; k0 < K; k0 += kRowThreads * kPer) {
        float hv[kPer];
        int kk[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int k = k0 + q * kRowThreads + sub;
            // a zero bin is skipped unless the slot is to be overwritten
            kk[q] = (k < K && (h[k] != 0.0f || pf_mode == kPfOverwrite)) ? k : -1;
            hv[q] = kk[q] >= 0 ? h[k] : 0.0f;
        }
        if (pf_mode == kPfOverwrite) {
#pragma unroll
            for (int q = 0; q < kPer; ++q)
                if (kk[q] >= 0) pflows[f0 + kk[q]] = (PF)hv[q];
        } else if (pf_mode == kPfAtomic) {
#pragma unroll
            for (int q = 0; q < kPer; ++q)
                if (kk[q] >= 0) atomicAdd(&pflows[f0 + kk[q]], (PF)hv[q]);
        } else {
            PF old[kPer];
#pragma unroll
            for (int q = 0; q < kPer; ++q) old[q] = kk[q] >= 0 ? pflows[f0 + kk[q]] : (PF)0;
#pragma unroll
            for (int q = 0; q < kPer; ++q)
                if (kk[q] >= 0) pflows[f0 + kk[q]] = old[q] + (PF)hv[q];
        }
    }
}

// Categorical flow backward without atomics per sample, for tiles of kSortTile nodes that
// share one variable and one category count K <= kSortK (the caller selects this kernel when
// every tile does; a tile that does not falls back to per-sample global atomics here).  Per batch chunk of kSortB columns the
// workgroup counting-sorts the variable's categories once (positions grouped by category), stages
// the tile's flow rows in LDS, and each thread sums, for its categories, the segment of every
// node.  The per-(node, category) sums stay in registers across chunks and are added to the
// parameter flows at the end (reads before writes, or atomics when slots are shared).
constexpr int kSortTile = 8;
constexpr int kSortK = 1024;
constexpr int kSortB = 1024;

template <typename PF>
__global__ __launch_bounds__(256) void cat_bwd_sorted_kernel(PF* __restrict__ pflows,
                                                             const float* __restrict__ node_flows,
                                                             const int* __restrict__ data, const int* __restrict__ vids,
                                                             const int64_t* __restrict__ s_pfids,
                                                             const int* __restrict__ ncats, int64_t num_nodes,
                                                             int64_t node_offset, int64_t B, int pf_mode) {
    __shared__ int cnt[kSortK + 1];
    __shared__ int cur[kSortK];
    __shared__ short perm[kSortB];
    __shared__ float fl[kSortTile][kSortB];
    __shared__ int uniform;
    constexpr int KPT = kSortK / 256;   // categories per thread
    const int tid = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * kSortTile;
    const int n_here = (int)((num_nodes - i0) < kSortTile ? (num_nodes - i0) : kSortTile);
    const int v = vids[i0];
    const int K = ncats[i0];
    if (tid == 0) {
        int ok = K <= kSortK;
        for (int r = 1; r < n_here; ++r) ok &= (vids[i0 + r] == v) && (ncats[i0 + r] == K);
        uniform = ok;
    }
    __syncthreads();
    if (!uniform) {
        if (pf_mode == kPfOverwrite) {
            for (int r = 0; r < n_here; ++r)
                for (int k = tid; k < ncats[i0 + r]; k += 256) pflows[s_pfids[i0 + r] + k] = (PF)0;
            __syncthreads();
        }
        for (int r = 0; r < n_here; ++r) {
            const int64_t i = i0 + r;
            const int* xr = data + (int64_t)vids[i] * B;
            const float* fr = node_flows + (node_offset + i) * B;
            for (int64_t b = tid; b < B; b += 256) {
                const int x = xr[b];
                if (x >= 0 && x < ncats[i] && fr[b] != 0.0f) atomicAdd(&pflows[s_pfids[i] + x], (PF)fr[b]);
            }
        }
        return;
    }
    const int* xrow = data + (int64_t)v * B;
    float sums[kSortTile][KPT];
#pragma unroll
    for (int r = 0; r < kSortTile; ++r)
#pragma unroll
        for (int q = 0; q < KPT; ++q) sums[r][q] = 0.0f;

    for (int64_t c0 = 0; c0 < B; c0 += kSortB) {
        const int nb = (int)((B - c0) < kSortB ? (B - c0) : kSortB);
        for (int k = tid; k <= K; k += 256) cnt[k] = 0;
        __syncthreads();
        // category counts and the tile's flow rows of this chunk
        for (int t = tid; t < nb; t += 256) {
            const int x = xrow[c0 + t];
            if (x >= 0 && x < K) atomicAdd(&cnt[x + 1], 1);
        }
        for (int t = tid; t < n_here * nb; t += 256) {
            const int r = t / nb, c = t - r * nb;
            fl[r][c] = node_flows[(node_offset + i0 + r) * B + c0 + c];
        }
        __syncthreads();
        // inclusive scan of cnt[0..K] by one wavefront (cnt[k] becomes the first position of
        // category k, since category x was counted at x + 1)
        if (tid < 64) {
            const int per = (K + 1 + 63) / 64;
            const int base = tid * per;
            int local = 0;
            for (int q = 0; q < per; ++q)
                if (base + q <= K) local += cnt[base + q];
            int incl = local;
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(incl, o, 64);
                if (tid >= o) incl += y;
            }
            int run = incl - local;
            for (int q = 0; q < per; ++q)
                if (base + q <= K) { run += cnt[base + q]; cnt[base + q] = run; }
        }
        __syncthreads();
        for (int k = tid; k < K; k += 256) cur[k] = cnt[k];
        __syncthreads();
        for (int t = tid; t < nb; t += 256) {
            const int x = xrow[c0 + t];
            if (x >= 0 && x < K) perm[atomicAdd(&cur[x], 1)] = (short)t;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            const int k = tid + q * 256;
            if (k < K) {
                const int p0 = cnt[k], p1 = cnt[k + 1];
                for (int p = p0; p < p1; ++p) {
                    const int c = perm[p];
#pragma unroll
                    for (int r = 0; r < kSortTile; ++r) sums[r][q] += fl[r][c];
                }
            }
        }
        __syncthreads();
    }
    // parameter flows: reads first, then writes (or plain stores when the accumulator is zero)
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
        const int k = tid + q * 256;
        if (k >= K) continue;
        if (pf_mode == kPfOverwrite) {
#pragma unroll
            for (int r = 0; r < kSortTile; ++r)
                if (r < n_here) pflows[s_pfids[i0 + r] + k] = (PF)sums[r][q];
        } else if (pf_mode == kPfAtomic) {
#pragma unroll
            for (int r = 0; r < kSortTile; ++r)
                if (r < n_here && sums[r][q] != 0.0f) atomicAdd(&pflows[s_pfids[i0 + r] + k], (PF)sums[r][q]);
        } else {
            PF old[kSortTile];
#pragma unroll
            for (int r = 0; r < kSortTile; ++r) old[r] = (r < n_here) ? pflows[s_pfids[i0 + r] + k] : (PF)0;
#pragma unroll
            for (int r = 0; r < kSortTile; ++r)
                if (r < n_here) pflows[s_pfids[i0 + r] + k] = old[r] + (PF)sums[r][q];
        }
    }
}

// The same sums when the whole batch fits one chunk (B <= SB), over tiles of TILE consecutive
// nodes (TILE divides kSortTile, so a tile lies inside one uniform tile of the caller's promise).
// Workgroup w takes the contiguous tile range [w T / G, (w + 1) T / G) of the T tiles (G
// workgroups: a whole multiple of as many as the chip keeps resident, so every slot gets the same
// share and no partial last wave of workgroups trails the launch).  A variable's counting sort is done once per
// workgroup for all its tiles of that variable (the tiles of one variable are consecutive), each
// tile's flow rows are staged with 16-byte loads, and each (node, category) sum is written once.  The flow
// staging takes TILE * SB * 4 bytes of LDS: 16 KiB (8 x 512 columns, 2 x 2048 columns) or 32 KiB
// (8 x 1024 columns); the smaller forms keep more workgroups resident.
//
// EM (pjamd_input_backward_em): the slots are stored, and each node's parameters are updated in
// place from them as pjamd_input_em would (the node's whole histogram is in the workgroup): the
// tile's parameters are loaded beside its flow rows, the flow sums reduced per node over the
// workgroup, and the new parameters stored.  Needs every node to own its slots and parameters.

// EM step of the fused backward (pjamd_input_backward_em); unused by the plain backward.
struct CatEm {
    float* params;
    const int64_t* s_pids;
    double step, pc;
    bool keep_zero;
};

// One tile's loads of cat_bwd_group_kernel, issued a tile ahead: its flow rows (thread share of
// NF float4) and, with EM, its parameters.  Branch-free: the caller passes a real tile (past its
// range's last tile, the range's first, never used); entries past the tile are clamped to real
// ones.
typedef float f32x4 __attribute__((ext_vector_type(4)));   // a register vector (float4 copies lower to memcpy)

template <int TILE, int NF, bool EM, int KPT>
__device__ __forceinline__ void cat_group_prefetch(f32x4 (&fr)[NF], float (&thn)[EM ? KPT : 1][EM ? TILE : 1],
                                                   const float* __restrict__ node_flows, const int* __restrict__ ncats,
                                                   int64_t p0, int64_t num_nodes, int64_t node_offset,
                                                   int64_t B, int nq, int tid, const CatEm& em) {
    const int nh = (int)((num_nodes - p0) < TILE ? (num_nodes - p0) : TILE);
#pragma unroll
    for (int u = 0; u < NF; ++u) {
        const int t = tid + 256 * u;
        const int tc = t < nh * nq ? t : 0;
        const int r = tc / nq, c = tc - r * nq;
        fr[u] = *reinterpret_cast<const f32x4*>(node_flows + (node_offset + p0 + r) * B + 4 * c);
    }
    if constexpr (EM) {
        const int kt = ncats[p0];
#pragma unroll
        for (int r = 0; r < TILE; ++r) {
            const int64_t pr = em.s_pids[p0 + (r < nh ? r : 0)];
#pragma unroll
            for (int q = 0; q < KPT; ++q) {
                const int k = tid + q * 256;
                thn[q][r] = em.params[pr + (k < kt ? k : 0)];
            }
        }
    }
}

// Waves per SIMD the register allocation aims at: 5 for the EM form with one category per thread
// (96 registers, no spills; the compiler's own choice, 102, leaves 4); no constraint otherwise.
template <typename PF, int TILE, int SB, bool EM = false, int KPT = kSortK / 256>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((EM && KPT == 1) ? 5 : 1, 8))) void cat_bwd_group_kernel(PF* __restrict__ pflows,
                                                            const float* __restrict__ node_flows,
                                                            const int* __restrict__ data, const int* __restrict__ vids,
                                                            const int64_t* __restrict__ s_pfids,
                                                            const int* __restrict__ ncats, int64_t num_nodes,
                                                            int64_t node_offset, int64_t B, int pf_mode,
                                                            CatEm em = CatEm{}) {
    __shared__ int cnt[kSortK + 1];
    __shared__ int cur[kSortK];
    __shared__ short perm[SB];
    __shared__ __attribute__((aligned(16))) float fl[TILE][SB];
    __shared__ double emred[EM ? 4 : 1][TILE];
    __shared__ double eminv[EM ? TILE : 1];
    static_assert(kSortTile % TILE == 0 && SB <= 32767, "tile and batch limits");
    static_assert(KPT >= 1 && KPT * 256 <= kSortK, "categories per thread");   // the caller keeps K <= 256 KPT
    const int tid = threadIdx.x;
    const int nb = (int)B;                                  // B <= SB
    const int nq = nb / 4;
    const int64_t ntiles = (num_nodes + TILE - 1) / TILE;
    const int64_t t0 = (int64_t)blockIdx.x * ntiles / gridDim.x;
    const int64_t t1 = ((int64_t)blockIdx.x + 1) * ntiles / gridDim.x;
    if (t0 >= t1) return;
    // Software pipeline: a tile's flow rows (and, with EM, its parameters) are loaded into
    // registers while the previous tile is summed and updated, then copied into LDS.
    constexpr int NF = TILE * SB / 1024;                    // float4 of a tile's flow rows per thread
    f32x4 fr[NF];
    float thn[EM ? KPT : 1][EM ? TILE : 1];
    cat_group_prefetch<TILE, NF, EM, KPT>(fr, thn, node_flows, ncats, t0 * TILE, num_nodes, node_offset, B, nq, tid,
                                          em);
    int cur_v = -1, K = 0;
    for (int64_t tt = t0; tt < t1; ++tt) {
        const int64_t i0 = tt * TILE;
        const int n_here = (int)((num_nodes - i0) < TILE ? (num_nodes - i0) : TILE);
        const int v = vids[i0];
        const int k_tile = ncats[i0];
        // Tiles of one variable may differ in category count (two Categorical vectors of
        // different num_cats over the same variable): the sort is redone when either changes.
        if (v != cur_v || k_tile != K) {
            // counting sort of the variable's categories: cnt[k] = first position of category k
            cur_v = v;
            K = k_tile;
            const int* xrow = data + (int64_t)v * B;
            for (int k = tid; k <= K; k += 256) cnt[k] = 0;
            __syncthreads();
            for (int t = tid; t < nb; t += 256) {
                const int x = xrow[t];
                if (x >= 0 && x < K) atomicAdd(&cnt[x + 1], 1);
            }
            __syncthreads();
            if (tid < 64) {
                const int per = (K + 1 + 63) / 64;
                const int base = tid * per;
                int local = 0;
                for (int q = 0; q < per; ++q)
                    if (base + q <= K) local += cnt[base + q];
                int incl = local;
                for (int o = 1; o < 64; o <<= 1) {
                    const int y = __shfl_up(incl, o, 64);
                    if (tid >= o) incl += y;
                }
                int run = incl - local;
                for (int q = 0; q < per; ++q)
                    if (base + q <= K) { run += cnt[base + q]; cnt[base + q] = run; }
            }
            __syncthreads();
            for (int k = tid; k < K; k += 256) cur[k] = cnt[k];
            __syncthreads();
            for (int t = tid; t < nb; t += 256) {
                const int x = xrow[t];
                if (x >= 0 && x < K) perm[atomicAdd(&cur[x], 1)] = (short)t;
            }
        }
        // the tile's flow rows (B % 4 == 0: 16-byte loads, issued one tile ahead)
#pragma unroll
        for (int u = 0; u < NF; ++u) {
            const int t = tid + 256 * u;
            if (t < n_here * nq) {
                const int r = t / nq, c = t - r * nq;
                *reinterpret_cast<f32x4*>(&fl[r][4 * c]) = fr[u];
            }
        }
        float th[EM ? KPT : 1][EM ? TILE : 1];
        if constexpr (EM) {
#pragma unroll
            for (int q = 0; q < KPT; ++q)
#pragma unroll
                for (int r = 0; r < TILE; ++r) th[q][r] = thn[q][r];
        }
        __syncthreads();
        // the staged registers are free: the next tile's loads go out before the sums, so they have
        // the sums, the stores and the EM step to land (headline EM form 287 -> 275 us, 2048
        // columns 1264 -> 1200 us, against issuing them after the sums)
        cat_group_prefetch<TILE, NF, EM, KPT>(fr, thn, node_flows, ncats, (tt + 1 < t1 ? tt + 1 : t0) * TILE,
                                              num_nodes, node_offset, B, nq, tid, em);
        float sq[EM ? KPT : 1][EM ? TILE : 1];   // EM: the slot values, kept for the update
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            const int k = tid + q * 256;
            if constexpr (EM) {
#pragma unroll
                for (int r = 0; r < TILE; ++r) sq[q][r] = 0.0f;
            }
            if (k >= K) continue;
            float sums[TILE];
#pragma unroll
            for (int r = 0; r < TILE; ++r) sums[r] = 0.0f;
            const int p0 = cnt[k], p1 = cnt[k + 1];
            for (int p = p0; p < p1; ++p) {
                const int c = perm[p];
#pragma unroll
                for (int r = 0; r < TILE; ++r) sums[r] += fl[r][c];
            }
            if constexpr (EM) {
#pragma unroll
                for (int r = 0; r < TILE; ++r) sq[q][r] = sums[r];
            }
            if (pf_mode == kPfOverwrite) {
#pragma unroll
                for (int r = 0; r < TILE; ++r)
                    if (r < n_here) pflows[s_pfids[i0 + r] + k] = (PF)sums[r];
            } else if (pf_mode == kPfAtomic) {
#pragma unroll
                for (int r = 0; r < TILE; ++r)
                    if (r < n_here && sums[r] != 0.0f) atomicAdd(&pflows[s_pfids[i0 + r] + k], (PF)sums[r]);
            } else {
                PF old[TILE];
#pragma unroll
                for (int r = 0; r < TILE; ++r) old[r] = (r < n_here) ? pflows[s_pfids[i0 + r] + k] : (PF)0;
#pragma unroll
                for (int r = 0; r < TILE; ++r)
                    if (r < n_here) pflows[s_pfids[i0 + r] + k] = old[r] + (PF)sums[r];
            }
        }
        if constexpr (EM) {
            // per node: the flow total (or, keeping zeros, the pseudocounted flows of the nonzero
            // parameters) over its K categories, as pjamd_input_em sums them, from the stored values
            const double npc = em.pc / (double)K;
            const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
            for (int r = 0; r < TILE; ++r) {
                double part = 0.0;
#pragma unroll
                for (int q = 0; q < KPT; ++q) {
                    if (tid + q * 256 >= K) continue;
                    const double f = (double)(PF)sq[q][r];
                    part += em.keep_zero ? ((th[q][r] < 1e-12f) ? 0.0 : f + npc) : f;
                }
                part = warp_sum(part);
                if (lane == 0) emred[wv][r] = part;
            }
            __syncthreads();
            // one division per node: step / denominator, shared through LDS
            if (tid < TILE) {
                const double tot = emred[0][tid] + emred[1][tid] + emred[2][tid] + emred[3][tid];
                eminv[tid] = em.step / (em.keep_zero ? tot : tot + em.pc);
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < TILE; ++r) {
                if (r >= n_here) continue;
                const double sc = eminv[r];
                const int64_t pr = em.s_pids[i0 + r];
#pragma unroll
                for (int q = 0; q < KPT; ++q) {
                    const int k = tid + q * 256;
                    if (k >= K) continue;
                    const double t = (double)th[q][r], f = (double)(PF)sq[q][r];
                    double nv = (1.0 - em.step) * t + (f + npc) * sc;
                    if (em.keep_zero && t < 1e-12) nv = 0.0;
                    em.params[pr + k] = (float)nv;
                }
            }
        }
        __syncthreads();
    }
}

// Clears the parameter-flow slots of every node (used before the atomic path when the caller
// asked for the flows to be overwritten).
template <typename PF>
__global__ __launch_bounds__(256) void cat_zero_slots_kernel(PF* __restrict__ pflows, const int64_t* __restrict__ s_pfids,
                                                             const int* __restrict__ ncats) {
    const int64_t i = blockIdx.x;
    for (int k = threadIdx.x; k < ncats[i]; k += blockDim.x) pflows[s_pfids[i] + k] = (PF)0;
}

// Categorical flow backward for very large category counts: direct global atomics.
template <typename PF>
__global__ __launch_bounds__(256) void cat_bwd_global_kernel(PF* __restrict__ pflows, const float* __restrict__ node_flows,
                                                             const int* __restrict__ data, const int* __restrict__ vids,
                                                             const int64_t* __restrict__ s_pfids,
                                                             const int* __restrict__ ncats, int64_t num_nodes,
                                                             int64_t node_offset, int64_t B) {
    const int64_t b = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.x;
    if (b >= B) return;
    const int x = data[(int64_t)vids[i] * B + b];
    const float f = node_flows[(node_offset + i) * B + b];
    if (x >= 0 && x < ncats[i] && f != 0.0f) atomicAdd(&pflows[s_pfids[i] + x], (PF)f);
}

// Categorical flow backward for tied vectors that share a flow range: one workgroup per range
// (group g = nodes gnodes[goff[g] .. goff[g + 1]), all of one category count K), whose four
// wavefronts take the members in turn and build one flow histogram over the range's K slots in
// LDS (K floats of dynamic LDS); then each slot is written once, reads before writes, zero bins
// skipped unless the slots are to be overwritten.  The group is the only writer of its range, so
// no global atomics are needed.  VEC: 16-byte loads (B % 4 == 0, 16-byte aligned rows).
constexpr int kTiedMaxK = 16384;

template <typename PF, bool VEC>
__global__ __launch_bounds__(256) void cat_bwd_tied_kernel(PF* __restrict__ pflows,
                                                           const float* __restrict__ node_flows,
                                                           const int* __restrict__ data, const int* __restrict__ vids,
                                                           const int64_t* __restrict__ s_pfids,
                                                           const int* __restrict__ ncats, const int* __restrict__ goff,
                                                           const int* __restrict__ gnodes, int64_t node_offset,
                                                           int64_t B, int pf_mode) {
    extern __shared__ float hist[];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int a = goff[blockIdx.x], e = goff[blockIdx.x + 1];
    const int i0 = gnodes[a];
    const int K = ncats[i0];
    for (int k = tid; k < K; k += 256) hist[k] = 0.0f;
    __syncthreads();
    for (int m = a + w; m < e; m += 4) {
        const int i = gnodes[m];
        const int* xrow = data + (int64_t)vids[i] * B;
        const float* frow = node_flows + (node_offset + i) * B;
        if constexpr (VEC) {
            for (int64_t b = (int64_t)lane * 4; b < B; b += 256) {
                const int4 x = *reinterpret_cast<const int4*>(xrow + b);
                const float4 f = *reinterpret_cast<const float4*>(frow + b);
                if (x.x >= 0 && x.x < K && f.x != 0.0f) atomicAdd(&hist[x.x], f.x);
                if (x.y >= 0 && x.y < K && f.y != 0.0f) atomicAdd(&hist[x.y], f.y);
                if (x.z >= 0 && x.z < K && f.z != 0.0f) atomicAdd(&hist[x.z], f.z);
                if (x.w >= 0 && x.w < K && f.w != 0.0f) atomicAdd(&hist[x.w], f.w);
            }
        } else {
            for (int64_t b = lane; b < B; b += 64) {
                const int x = xrow[b];
                const float f = frow[b];
                if (x >= 0 && x < K && f != 0.0f) atomicAdd(&hist[x], f);
            }
        }
    }
    __syncthreads();
    constexpr int kPer = 8;
    const int64_t f0 = s_pfids[i0];
    for (int k0 = 0; k0 < K; k0 += 256 * kPer) {
        float hv[kPer];
        int kk[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int k = k0 + q * 256 + tid;
            kk[q] = (k < K && (pf_mode == kPfOverwrite || hist[k] != 0.0f)) ? k : -1;
            hv[q] = kk[q] >= 0 ? hist[k] : 0.0f;
        }
        if (pf_mode == kPfOverwrite) {
#pragma unroll
            for (int q = 0; q < kPer; ++q)
                if (kk[q] >= 0) pflows[f0 + kk[q]] = (PF)hv[q];
        } else {
            PF old[kPer];
#pragma unroll
            for (int q = 0; q < kPer; ++q) old[q] = kk[q] >= 0 ? pflows[f0 + kk[q]] : (PF)0;
#pragma unroll
            for (int q = 0; q < kPer; ++q)
                if (kk[q] >= 0) pflows[f0 + kk[q]] = old[q] + (PF)hv[q];
        }
    }
}

// Bernoulli flow backward: slot 0 collects flows of x = 1, slot 1 of x = 0.
template <typename PF>
__global__ __launch_bounds__(256) void bern_bwd_kernel(PF* __restrict__ pflows, const float* __restrict__ node_flows,
                                                       const int* __restrict__ data, const int* __restrict__ vids,
                                                       const int64_t* __restrict__ s_pfids, int64_t node_offset,
                                                       int64_t B, int atomic_out) {
    __shared__ float red[2][4];
    const int64_t i = blockIdx.x;
    float s1 = 0.0f, s0 = 0.0f;
    for (int64_t b = threadIdx.x; b < B; b += blockDim.x) {
        const int x = data[(int64_t)vids[i] * B + b];
        const float f = node_flows[(node_offset + i) * B + b];
        if (x > 0) s1 += f; else if (x == 0) s0 += f;     // x < 0: marginalized, no flow
    }
    s1 = warp_sum(s1);
    s0 = warp_sum(s0);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { red[0][w] = s1; red[1][w] = s0; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = 0.0f, c = 0.0f;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) { a += red[0][q]; c += red[1][q]; }
        pf_add<PF>(&pflows[s_pfids[i]], (PF)a, atomic_out);
        pf_add<PF>(&pflows[s_pfids[i] + 1], (PF)c, atomic_out);
    }
}

// Flows of marginalized entries (negative categories), Categorical: a node's flow on a missing
// entry is spread over its categories in proportion to theta (the posterior of the category when
// the variable is summed out): param_flows[s_pfid_i + k] += theta_ik * sum of those flows.  One
// workgroup per node; runs after the per-category pass, which skips these entries.
template <typename PF>
__global__ __launch_bounds__(256) void cat_missing_flows_kernel(PF* __restrict__ pflows,
                                                                const float* __restrict__ node_flows,
                                                                const float* __restrict__ params,
                                                                const int* __restrict__ data,
                                                                const int* __restrict__ vids,
                                                                const int64_t* __restrict__ s_pids,
                                                                const int64_t* __restrict__ s_pfids,
                                                                const int* __restrict__ ncats, int64_t node_offset,
                                                                int64_t B) {
    __shared__ float red[4];
    const int64_t i = blockIdx.x;
    const int* xr = data + (int64_t)vids[i] * B;
    const float* fr = node_flows + (node_offset + i) * B;
    float f = 0.0f;
    for (int64_t b = threadIdx.x; b < B; b += blockDim.x) f += xr[b] < 0 ? fr[b] : 0.0f;
    f = warp_sum(f);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) red[w] = f;
    __syncthreads();
    const float tot = red[0] + red[1] + red[2] + red[3];
    if (tot == 0.0f) return;
    const float* th = params + s_pids[i];
    for (int k = threadIdx.x; k < ncats[i]; k += blockDim.x)
        atomicAdd(&pflows[s_pfids[i] + k], (PF)th[k] * (PF)tot);
}

// Categorical EM, one wavefront per source node; sums and the update run in double.  Nodes of at
// most 64 * KR categories (KR = 4: 256, KR = 16: 1024) hold their flows and parameters in
// registers: every lane issues its KR loads of each at once (index clamped, value masked on use),
// so the normaliser and the update cost one memory round trip; wider nodes loop.
template <typename PF, int KR>
__device__ __forceinline__ void cat_em_regs(float* __restrict__ params, const PF* __restrict__ pflows, int K,
                                            int64_t p0, int64_t f0, int lane, double step, double pc,
                                            bool keep_zero) {
    const double npc = pc / (double)K;
    double f[KR];
    float t[KR];
#pragma unroll
    for (int q = 0; q < KR; ++q) {
        const int k = lane + 64 * q;
        const int kc = k < K ? k : 0;
        f[q] = (double)pflows[f0 + kc];
        t[q] = params[p0 + kc];
    }
    double cum = 0.0;
#pragma unroll
    for (int q = 0; q < KR; ++q) {
        if (lane + 64 * q < K) cum += keep_zero ? ((t[q] < 1e-12f) ? 0.0 : f[q] + npc) : f[q];
    }
    cum = warp_sum(cum) + pc;
#pragma unroll
    for (int q = 0; q < KR; ++q) {
        const int k = lane + 64 * q;
        if (k >= K) continue;
        const double th = (double)t[q];
        double nv;
        if (keep_zero) nv = (th < 1e-12) ? 0.0 : (1.0 - step) * th + step * (f[q] + npc) / (cum - pc);
        else nv = (1.0 - step) * th + step * (f[q] + npc) / cum;
        params[p0 + k] = (float)nv;
    }
}

template <typename PF>
__global__ __launch_bounds__(256) void cat_em_kernel(float* __restrict__ params, const PF* __restrict__ pflows,
                                                     const int* __restrict__ src, const int64_t* __restrict__ s_pids,
                                                     const int64_t* __restrict__ s_pfids, const int* __restrict__ ncats,
                                                     int64_t num_src, double step, double pc, bool keep_zero) {
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= num_src) return;
    const int i = src[w];
    const int K = ncats[i];
    const int64_t p0 = s_pids[i], f0 = s_pfids[i];
    if (K <= 256) {
        cat_em_regs<PF, 4>(params, pflows, K, p0, f0, lane, step, pc, keep_zero);
        return;
    }
    if (K <= 1024) {
        cat_em_regs<PF, 16>(params, pflows, K, p0, f0, lane, step, pc, keep_zero);
        return;
    }
    const double npc = pc / (double)K;
    double cum = 0.0;
    for (int k = lane; k < K; k += 64) {
        const double f = (double)pflows[f0 + k];
        if (keep_zero) {
            cum += (params[p0 + k] < 1e-12f) ? 0.0 : f + npc;
        } else {
            cum += f;
        }
    }
    cum = warp_sum(cum) + pc;
    for (int k = lane; k < K; k += 64) {
        const double th = (double)params[p0 + k];
        const double f = (double)pflows[f0 + k];
        double nv;
        if (keep_zero) {
            nv = (th < 1e-12) ? 0.0 : (1.0 - step) * th + step * (f + npc) / (cum - pc);
        } else {
            nv = (1.0 - step) * th + step * (f + npc) / cum;
        }
        params[p0 + k] = (float)nv;
    }
}

template <typename PF>
__global__ __launch_bounds__(256) void bern_em_kernel(float* __restrict__ params, const PF* __restrict__ pflows,
                                                      const int* __restrict__ src, const int64_t* __restrict__ s_pids,
                                                      const int64_t* __restrict__ s_pfids, int64_t num_src,
                                                      double step, double pc) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= num_src) return;
    const int i = src[w];
    const double f1 = (double)pflows[s_pfids[i]], f0 = (double)pflows[s_pfids[i] + 1];
    const double th = (double)params[s_pids[i]];
    const double upd = (f1 + 0.5 * pc) / (f1 + f0 + pc);
    params[s_pids[i]] = (float)((1.0 - step) * th + step * upd);
}

// Input sampling.  A workgroup takes one node and 256 samples; if any of them is active (flags
// cleared as they are read), a Categorical node's parameters are staged in LDS, one thread turns
// them into FP32 running sums (left to right), and every active sample takes the first category
// whose running sum reaches u (binary search; 0 if none).  Category counts above kCatLds scan the
// parameters per sample instead, kCatBatch loads at a time.
constexpr int kCatLds = 4096;
constexpr int kCatBatch = 16;

template <typename OUT>
__global__ __launch_bounds__(256) void input_sample_kernel(int kind, OUT* __restrict__ out, const float* __restrict__ params,
                                                           const float* __restrict__ cdf, uint8_t* __restrict__ node_act,
                                                           const float* __restrict__ uniforms,
                                                           const uint32_t* __restrict__ seed, const int* __restrict__ vids,
                                                           const int64_t* __restrict__ s_pids, const int* __restrict__ ncats,
                                                           int64_t num_nodes, int64_t node_offset, int64_t S) {
    __shared__ float P[kCatLds];
    const int tid = threadIdx.x;
    const int64_t s = (int64_t)blockIdx.y * blockDim.x + tid;
    const int64_t i = blockIdx.x;
    const int64_t row = (node_offset + i) * S + s;
    const bool act = s < S && node_act[row] != 0;
    if (!__syncthreads_or(act)) return;
    float u = 0.0f;
    if (act) {
        node_act[row] = 0;
        u = uniforms != nullptr ? uniforms[row] : philox_uniform(seed[0], seed[1], (uint32_t)s, (uint32_t)(node_offset + i));
    }
    const float* th = params + s_pids[i];
    int pick = 0;
    if (kind == 1) {
        pick = (u < th[0]) ? 1 : 0;
    } else {
        const int K = ncats[i];
        if (K <= kCatLds) {
            // the parameters, or their running sums when precomputed (pjamd_sample_prefix)
            const float* src = cdf != nullptr ? cdf + s_pids[i] : th;
            for (int k = tid; k < K; k += blockDim.x) P[k] = src[k];
            __syncthreads();
            if (tid == 0 && cdf == nullptr) {
                float c = 0.0f;
                for (int k = 0; k < K; ++k) {
                    c += P[k];
                    P[k] = c;
                }
            }
            __syncthreads();
            if (act && K > 0 && P[K - 1] >= u) {
                int lo = 0, hi = K - 1;        // first k with P[k] >= u
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (P[mid] >= u) hi = mid;
                    else lo = mid + 1;
                }
                pick = lo;
            }
        } else if (act) {
            float cum = 0.0f;
            int found = -1;
            for (int k0 = 0; k0 < K && found < 0; k0 += kCatBatch) {
                float w[kCatBatch];
#pragma unroll
                for (int q = 0; q < kCatBatch; ++q) w[q] = th[min(k0 + q, K - 1)];
#pragma unroll
                for (int q = 0; q < kCatBatch; ++q) {
                    if (k0 + q < K && found < 0) {
                        cum += w[q];
                        if (cum >= u) found = k0 + q;
                    }
                }
            }
            pick = found < 0 ? 0 : found;
        }
    }
    if (act) out[(int64_t)vids[i] * S + s] = (OUT)pick;
}

}  // namespace

PJAMD_EXPORT int pjamd_input_forward(int kind, const float* params, float* node_mars, const int* data,
                                     const int* vids, const int64_t* s_pids, const int* ncats, int64_t num_nodes,
                                     int64_t node_offset, int64_t B, int kmax, hipStream_t stream) {
    PJAMD_REQUIRE(num_nodes >= 0 && B > 0 && (kind == 0 || kind == 1));
    if (num_nodes == 0) return PJAMD_OK;
    PJAMD_REQUIRE(kind == 1 || kmax > 0);
    // 16-byte rows: B a multiple of 4 and 16-byte aligned buffers (int32 data rows, float rows)
    const bool vec = (B % 4 == 0) && ((((uintptr_t)data) | ((uintptr_t)node_mars)) % 16 == 0) &&
                     ((node_offset * B) % 4 == 0);
    if (!vec) {
        dim3 grid((unsigned)num_nodes, cdiv(B, 256));
        cat_fwd_global_kernel<<<grid, 256, 0, stream>>>(params, node_mars, data, vids, s_pids, ncats, node_offset, B,
                                                        kind);
        PJAMD_LAUNCH_CHECK();
        return PJAMD_OK;
    }
    if (kind == 0 && kmax <= kFwdLdsK) {
        const int64_t wblocks = cdiv(num_nodes, (int64_t)4 * kFwdNodes);
        PJAMD_REQUIRE(wblocks < ((int64_t)1 << 31));
        input_fwd_lds_kernel<kFwdNodes><<<(unsigned)wblocks, 256, 0, stream>>>(params, node_mars, data, vids, s_pids,
                                                                               ncats, num_nodes, node_offset, B);
        PJAMD_LAUNCH_CHECK();
        return PJAMD_OK;
    }
    const int64_t blocks = cdiv(num_nodes, (int64_t)(256 / kFwdRowThreads) * kFwdNodes);
    PJAMD_REQUIRE(blocks < ((int64_t)1 << 31));
    if (kind == 0)
        input_fwd_kernel<0, kFwdNodes><<<(unsigned)blocks, 256, 0, stream>>>(params, node_mars, data, vids, s_pids,
                                                                             ncats, num_nodes, node_offset, B);
    else
        input_fwd_kernel<1, kFwdNodes><<<(unsigned)blocks, 256, 0, stream>>>(params, node_mars, data, vids, s_pids,
                                                                             ncats, num_nodes, node_offset, B);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

// Launches cat_bwd_group_kernel over ceil(num_nodes / TILE) tiles with min(tiles, kPerSlot x
// resident) workgroups, resident = the kernel's occupancy per CU (looked up once per instance) x
// the CUs.
template <typename PF, int TILE, int SB, bool EM = false, int KPT = kSortK / 256>
static int launch_group(PF* pflows, const float* node_flows, const int* data, const int* vids, const int64_t* s_pfids,
                        const int* ncats, int64_t num_nodes, int64_t node_offset, int64_t B, int pf_mode,
                        const CatEm& em, hipStream_t stream) {
    static std::atomic<int> occupancy{0};
    int occ = occupancy.load(std::memory_order_relaxed);
    hipError_t e;
    if (occ == 0) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, cat_bwd_group_kernel<PF, TILE, SB, EM, KPT>, 256, 0);
        if (e != hipSuccess) return (int)e;
        occ = occ > 0 ? occ : 1;
        occupancy.store(occ, std::memory_order_relaxed);
    }
    int dev = 0, cus = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return (int)e;
    // Workgroups per resident slot.  Up to 1024 columns: one (balanced ranges; the EM form at
    // HCLT-784 L=256, 512 columns: 287 us, against 293 / 299 us with 4 / 8; PD 28x28 L=256, 1024
    // columns: 0.42 ms, against 0.46 ms with 8).  At 2048 columns a tile's time follows its
    // variable's category skew (one thread sums a category's samples), so static ranges leave the
    // heavy ones trailing and 8 per slot let the dispatcher even them out (HCLT-784 L=512: 1466 /
    // 1324 / 1281 / 1266 us with 1 / 2 / 4 / 8).
    constexpr int kPerSlot = SB <= 1024 ? 1 : 8;
    const int64_t tiles = cdiv(num_nodes, (int64_t)TILE), workers = (int64_t)occ * (cus > 0 ? cus : 1) * kPerSlot;
    const unsigned grid = (unsigned)(tiles < workers ? tiles : workers);
    cat_bwd_group_kernel<PF, TILE, SB, EM, KPT><<<grid, 256, 0, stream>>>(pflows, node_flows, data, vids, s_pfids, ncats,
                                                                          num_nodes, node_offset, B, pf_mode, em);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

template <typename PF>
static int input_backward_impl(int kind, PF* pflows, const float* node_flows, const int* data, const int* vids,
                               const int64_t* s_pfids, const int* ncats, int64_t num_nodes, int64_t node_offset,
                               int64_t B, int pf_mode, bool sorted_tiles, int kmax, hipStream_t stream) {
    if (kind == 1) {
        bern_bwd_kernel<PF><<<(unsigned)num_nodes, 256, 0, stream>>>(pflows, node_flows, data, vids, s_pfids,
                                                                    node_offset, B, pf_mode);
    } else if (sorted_tiles && kmax <= kSortK && B <= 2048 && B % 4 == 0 &&
               ((((uintptr_t)node_flows) | ((uintptr_t)data)) % 16 == 0) && (node_offset * B) % 4 == 0) {
        // 512 columns: 16 KiB of staged flows, twice the resident workgroups of the 1024-column form
        if (B <= 512)
            return launch_group<PF, 8, 512>(pflows, node_flows, data, vids, s_pfids, ncats, num_nodes, node_offset, B,
                                            pf_mode, CatEm{}, stream);
        if (B <= 1024)
            return launch_group<PF, 8, 1024>(pflows, node_flows, data, vids, s_pfids, ncats, num_nodes, node_offset, B,
                                             pf_mode, CatEm{}, stream);
        return launch_group<PF, 2, 2048>(pflows, node_flows, data, vids, s_pfids, ncats, num_nodes, node_offset, B,
                                         pf_mode, CatEm{}, stream);
    } else if (sorted_tiles && kmax <= kSortK) {
        cat_bwd_sorted_kernel<PF><<<(unsigned)cdiv(num_nodes, kSortTile), 256, 0, stream>>>(
            pflows, node_flows, data, vids, s_pfids, ncats, num_nodes, node_offset, B, pf_mode);
    } else if (kmax <= kLdsFloats && B % 4 == 0 && ((((uintptr_t)data) | ((uintptr_t)node_flows)) % 16 == 0)) {
        constexpr int NR = 256 / kRowThreads;
        const size_t lds = (size_t)NR * kmax * sizeof(float);
        cat_bwd_lds_kernel<PF><<<(unsigned)cdiv(num_nodes, NR), 256, lds, stream>>>(
            pflows, node_flows, data, vids, s_pfids, ncats, num_nodes, node_offset, B, kmax, pf_mode);
    } else {
        if (pf_mode == kPfOverwrite)
            cat_zero_slots_kernel<PF><<<(unsigned)num_nodes, 256, 0, stream>>>(pflows, s_pfids, ncats);
        dim3 grid((unsigned)num_nodes, cdiv(B, 256));
        cat_bwd_global_kernel<PF><<<grid, 256, 0, stream>>>(pflows, node_flows, data, vids, s_pfids, ncats, num_nodes,
                                                           node_offset, B);
    }
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_input_missing_flows(int kind, void* pflows, int pf64, const float* node_flows,
                                           const float* params, const int* data, const int* vids,
                                           const int64_t* s_pids, const int64_t* s_pfids, const int* ncats,
                                           int64_t num_nodes, int64_t node_offset, int64_t B, hipStream_t stream) {
    PJAMD_REQUIRE(num_nodes >= 0 && B > 0 && (kind == 0 || kind == 1) && num_nodes < ((int64_t)1 << 31));
    if (num_nodes == 0 || kind == 1) return PJAMD_OK;   // Bernoulli: missing entries carry no flow
    if (pf64)
        cat_missing_flows_kernel<double><<<(unsigned)num_nodes, 256, 0, stream>>>(
            (double*)pflows, node_flows, params, data, vids, s_pids, s_pfids, ncats, node_offset, B);
    else
        cat_missing_flows_kernel<float><<<(unsigned)num_nodes, 256, 0, stream>>>(
            (float*)pflows, node_flows, params, data, vids, s_pids, s_pfids, ncats, node_offset, B);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

template <typename PF>
static int input_backward_tied_impl(PF* pflows, const float* node_flows, const int* data, const int* vids,
                                    const int64_t* s_pfids, const int* ncats, const int* goff, const int* gnodes,
                                    int64_t n_groups, int kmax, int64_t node_offset, int64_t B, int pf_mode,
                                    hipStream_t stream) {
    const size_t lds = (size_t)kmax * sizeof(float);
    const bool vec = B % 4 == 0 && (node_offset * B) % 4 == 0 &&
                     ((((uintptr_t)node_flows) | ((uintptr_t)data)) % 16 == 0);
    if (vec)
        cat_bwd_tied_kernel<PF, true><<<(unsigned)n_groups, 256, lds, stream>>>(
            pflows, node_flows, data, vids, s_pfids, ncats, goff, gnodes, node_offset, B, pf_mode);
    else
        cat_bwd_tied_kernel<PF, false><<<(unsigned)n_groups, 256, lds, stream>>>(
            pflows, node_flows, data, vids, s_pfids, ncats, goff, gnodes, node_offset, B, pf_mode);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_input_backward_tied(int kind, void* pflows, const float* node_flows, int pf64, const int* data,
                                           const int* vids, const int64_t* s_pfids, const int* ncats,
                                           const int* group_offsets, const int* group_nodes, int64_t n_groups,
                                           int kmax, int64_t node_offset, int64_t B, int overwrite,
                                           hipStream_t stream) {
    PJAMD_REQUIRE(n_groups >= 0 && B > 0 && n_groups < ((int64_t)1 << 31) && kmax > 0);
    if (kind != 0 || kmax > kTiedMaxK) return PJAMD_EUNSUPPORTED;
    if (n_groups == 0) return PJAMD_OK;
    const int pf_mode = overwrite ? kPfOverwrite : kPfAdd;
    if (pf64)
        return input_backward_tied_impl<double>((double*)pflows, node_flows, data, vids, s_pfids, ncats, group_offsets,
                                                group_nodes, n_groups, kmax, node_offset, B, pf_mode, stream);
    return input_backward_tied_impl<float>((float*)pflows, node_flows, data, vids, s_pfids, ncats, group_offsets,
                                           group_nodes, n_groups, kmax, node_offset, B, pf_mode, stream);
}

// `layout`: bit 0 = some flow ranges are shared by several input vectors (atomic adds); bit 1 =
// every tile of kSortTile consecutive nodes shares one variable and category count; bit 2 = the
// accumulator is zero, store instead of add (not with bit 0); bits 3.. = largest category count
// of the layer (Categorical).
PJAMD_EXPORT int pjamd_input_backward(int kind, void* pflows, const float* node_flows, int pf64, const int* data,
                                      const int* vids, const int64_t* s_pfids, const int* ncats, int64_t num_nodes,
                                      int64_t node_offset, int64_t B, int layout, hipStream_t stream) {
    PJAMD_REQUIRE(num_nodes >= 0 && B > 0 && (kind == 0 || kind == 1) && layout >= 0);
    if (num_nodes == 0) return PJAMD_OK;
    PJAMD_REQUIRE(!((layout & 1) && (layout & 4)));
    const int pf_mode = (layout & 1) ? kPfAtomic : (layout & 4) ? kPfOverwrite : kPfAdd;
    const bool sorted_tiles = (layout & 2) != 0;
    const int kmax = layout >> 3;
    PJAMD_REQUIRE(kind == 1 || kmax > 0);
    if (pf64)
        return input_backward_impl<double>(kind, (double*)pflows, node_flows, data, vids, s_pfids, ncats, num_nodes,
                                           node_offset, B, pf_mode, sorted_tiles, kmax, stream);
    return input_backward_impl<float>(kind, (float*)pflows, node_flows, data, vids, s_pfids, ncats, num_nodes,
                                      node_offset, B, pf_mode, sorted_tiles, kmax, stream);
}

template <typename PF, int KPT>
static int input_backward_em_impl(PF* pflows, const float* node_flows, const int* data, const int* vids,
                                  const int64_t* s_pfids, const int* ncats, int64_t num_nodes, int64_t node_offset,
                                  int64_t B, const CatEm& em, hipStream_t stream) {
    if (B <= 512)
        return launch_group<PF, 8, 512, true, KPT>(pflows, node_flows, data, vids, s_pfids, ncats, num_nodes,
                                                   node_offset, B, kPfOverwrite, em, stream);
    if (B <= 1024)
        return launch_group<PF, 8, 1024, true, KPT>(pflows, node_flows, data, vids, s_pfids, ncats, num_nodes,
                                                    node_offset, B, kPfOverwrite, em, stream);
    return launch_group<PF, 2, 2048, true, KPT>(pflows, node_flows, data, vids, s_pfids, ncats, num_nodes, node_offset,
                                                B, kPfOverwrite, em, stream);
}

PJAMD_EXPORT int pjamd_input_backward_em(void* pflows, const float* node_flows, int pf64, const int* data,
                                         const int* vids, const int64_t* s_pfids, const int* ncats,
                                         int64_t num_nodes, int64_t node_offset, int64_t B, int kmax, float* params,
                                         const int64_t* s_pids, float step, float pseudocount, int keep_zero,
                                         hipStream_t stream) {
    PJAMD_REQUIRE(num_nodes >= 0 && B > 0 && kmax > 0 && step > 0.0f && step <= 1.0f);
    if (num_nodes == 0) return PJAMD_OK;
    PJAMD_REQUIRE(num_nodes < ((int64_t)1 << 31) && params && s_pids);
    const bool vec = B % 4 == 0 && (node_offset * B) % 4 == 0 &&
                     ((((uintptr_t)node_flows) | ((uintptr_t)data)) % 16 == 0);
    if (!(vec && B <= 2048 && kmax <= kSortK)) return PJAMD_EUNSUPPORTED;
    const CatEm em{params, s_pids, (double)step, (double)pseudocount, keep_zero != 0};
    // categories per thread: one up to 256 (K <= 256 KPT)
    if (pf64) {
        if (kmax <= 256)
            return input_backward_em_impl<double, 1>((double*)pflows, node_flows, data, vids, s_pfids, ncats,
                                                     num_nodes, node_offset, B, em, stream);
        return input_backward_em_impl<double, 4>((double*)pflows, node_flows, data, vids, s_pfids, ncats, num_nodes,
                                                 node_offset, B, em, stream);
    }
    if (kmax <= 256)
        return input_backward_em_impl<float, 1>((float*)pflows, node_flows, data, vids, s_pfids, ncats, num_nodes,
                                                node_offset, B, em, stream);
    return input_backward_em_impl<float, 4>((float*)pflows, node_flows, data, vids, s_pfids, ncats, num_nodes,
                                            node_offset, B, em, stream);
}

PJAMD_EXPORT int pjamd_input_em(int kind, float* params, const void* pflows, int pf64, const int* src,
                                const int64_t* s_pids, const int64_t* s_pfids, const int* ncats, int64_t num_src,
                                float step, float pseudocount, int keep_zero, hipStream_t stream) {
    PJAMD_REQUIRE(num_src >= 0 && (kind == 0 || kind == 1));
    if (num_src == 0) return PJAMD_OK;
    if (kind == 0) {
        if (pf64)
            cat_em_kernel<double><<<cdiv(num_src, 4), 256, 0, stream>>>(params, (const double*)pflows, src, s_pids,
                                                                     s_pfids, ncats, num_src, step, pseudocount, keep_zero);
        else
            cat_em_kernel<float><<<cdiv(num_src, 4), 256, 0, stream>>>(params, (const float*)pflows, src, s_pids,
                                                                    s_pfids, ncats, num_src, step, pseudocount, keep_zero);
    } else {
        if (pf64)
            bern_em_kernel<double><<<cdiv(num_src, 256), 256, 0, stream>>>(params, (const double*)pflows, src, s_pids,
                                                                        s_pfids, num_src, step, pseudocount);
        else
            bern_em_kernel<float><<<cdiv(num_src, 256), 256, 0, stream>>>(params, (const float*)pflows, src, s_pids,
                                                                       s_pfids, num_src, step, pseudocount);
    }
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_input_sample(int kind, void* out, const float* params, const float* cdf, uint8_t* node_act,
                                    const float* uniforms, const uint32_t* seed, const int* vids,
                                    const int64_t* s_pids, const int* ncats, int out_i64, int64_t num_nodes,
                                    int64_t node_offset, int64_t S, hipStream_t stream) {
    PJAMD_REQUIRE(num_nodes >= 0 && S > 0 && (kind == 0 || kind == 1) && (uniforms != nullptr || seed != nullptr));
    if (num_nodes == 0) return PJAMD_OK;
    dim3 grid((unsigned)num_nodes, cdiv(S, 256));
    if (out_i64)
        input_sample_kernel<int64_t><<<grid, 256, 0, stream>>>(kind, (int64_t*)out, params, cdf, node_act, uniforms, seed,
                                                              vids, s_pids, ncats, num_nodes, node_offset, S);
    else
        input_sample_kernel<uint8_t><<<grid, 256, 0, stream>>>(kind, (uint8_t*)out, params, cdf, node_act, uniforms, seed,
                                                              vids, s_pids, ncats, num_nodes, node_offset, S);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}
