// Input layers: Categorical (kind 0) and Bernoulli (kind 1) forward, flow backward,
// mini-batch EM and sampling.  Data is [num_vars, B] int32; node rows are [num_nodes, B].
#include "pjamd_common.h"

using namespace pjamd;

namespace {

constexpr int kFwdRows = 4;  // input nodes per thread in the forward pass

// node_mars[(off + i) * B + b] = log p_i(x[v_i, b]); out-of-range categories have p = 0.
template <int KIND>
__global__ __launch_bounds__(256) void input_fwd_kernel(const float* __restrict__ params, float* __restrict__ node_mars,
                                                        const int* __restrict__ data, const int* __restrict__ vids,
                                                        const int64_t* __restrict__ s_pids,
                                                        const int* __restrict__ ncats, int64_t num_nodes,
                                                        int64_t node_offset, int64_t B) {
    const int64_t b = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int64_t i0 = (int64_t)blockIdx.x * kFwdRows;
    int last_v = -1;
    int x = 0;
#pragma unroll
    for (int r = 0; r < kFwdRows; ++r) {
        const int64_t i = i0 + r;
        if (i >= num_nodes) break;
        const int v = vids[i];
        if (v != last_v) {
            x = data[(int64_t)v * B + b];
            last_v = v;
        }
        float lp;
        if (KIND == 0) {
            const int K = ncats[i];
            lp = (x >= 0 && x < K) ? __logf(params[s_pids[i] + x]) : kNegInf;
        } else {
            const float p = params[s_pids[i]];
            lp = __logf(x > 0 ? p : 1.0f - p);
        }
        node_mars[(node_offset + i) * B + b] = lp;
    }
}

// Categorical flow backward: each workgroup owns `nt` consecutive input nodes and builds
// their [nt, kmax] flow histograms in LDS, then adds them to the parameter flows.
template <typename PF>
__global__ __launch_bounds__(256) void cat_bwd_lds_kernel(PF* __restrict__ pflows, const float* __restrict__ node_flows,
                                                          const int* __restrict__ data, const int* __restrict__ vids,
                                                          const int64_t* __restrict__ s_pfids,
                                                          const int* __restrict__ ncats, int64_t num_nodes,
                                                          int64_t node_offset, int64_t B, int nt, int kmax,
                                                          bool atomic_out) {
    extern __shared__ float hist[];
    const int64_t i0 = (int64_t)blockIdx.x * nt;
    const int n_here = (int)((num_nodes - i0) < nt ? (num_nodes - i0) : nt);
    for (int t = threadIdx.x; t < nt * kmax; t += blockDim.x) hist[t] = 0.0f;
    __syncthreads();
    for (int r = 0; r < n_here; ++r) {
        const int64_t i = i0 + r;
        const int* xrow = data + (int64_t)vids[i] * B;
        const float* frow = node_flows + (node_offset + i) * B;
        const int K = ncats[i];
        for (int64_t b = threadIdx.x; b < B; b += blockDim.x) {
            const int x = xrow[b];
            const float f = frow[b];
            if (x >= 0 && x < K && f != 0.0f) atomicAdd(&hist[r * kmax + x], f);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n_here * kmax; t += blockDim.x) {
        const int r = t / kmax;
        const int k = t - r * kmax;
        const int64_t i = i0 + r;
        if (k < ncats[i]) {
            const float h = hist[t];
            if (h != 0.0f) pf_add<PF>(&pflows[s_pfids[i] + k], (PF)h, atomic_out);
        }
    }
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

// Bernoulli flow backward: slot 0 collects flows of x = 1, slot 1 of x = 0.
template <typename PF>
__global__ __launch_bounds__(256) void bern_bwd_kernel(PF* __restrict__ pflows, const float* __restrict__ node_flows,
                                                       const int* __restrict__ data, const int* __restrict__ vids,
                                                       const int64_t* __restrict__ s_pfids, int64_t node_offset,
                                                       int64_t B, bool atomic_out) {
    __shared__ float red[2][4];
    const int64_t i = blockIdx.x;
    float s1 = 0.0f, s0 = 0.0f;
    for (int64_t b = threadIdx.x; b < B; b += blockDim.x) {
        const int x = data[(int64_t)vids[i] * B + b];
        const float f = node_flows[(node_offset + i) * B + b];
        if (x > 0) s1 += f; else s0 += f;
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

// Categorical EM, one wavefront per source node; sums and the update run in double.
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

template <typename OUT>
__global__ __launch_bounds__(256) void input_sample_kernel(int kind, OUT* __restrict__ out, const float* __restrict__ params,
                                                           const uint8_t* __restrict__ node_act,
                                                           const float* __restrict__ uniforms, const int* __restrict__ vids,
                                                           const int64_t* __restrict__ s_pids, const int* __restrict__ ncats,
                                                           int64_t num_nodes, int64_t node_offset, int64_t S) {
    const int64_t s = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.x;
    if (s >= S) return;
    const int64_t row = (node_offset + i) * S + s;
    if (!node_act[row]) return;
    const float u = uniforms[row];
    int pick = 0;
    if (kind == 0) {
        const int K = ncats[i];
        float cum = 0.0f;
        for (int k = 0; k < K; ++k) {
            cum += params[s_pids[i] + k];
            if (cum >= u) { pick = k; break; }
        }
    } else {
        pick = (u < params[s_pids[i]]) ? 1 : 0;
    }
    out[(int64_t)vids[i] * S + s] = (OUT)pick;
}

}  // namespace

PJAMD_EXPORT int pjamd_input_forward(int kind, const float* params, float* node_mars, const int* data,
                                     const int* vids, const int64_t* s_pids, const int* ncats, int64_t num_nodes,
                                     int64_t node_offset, int64_t B, hipStream_t stream) {
    PJAMD_REQUIRE(num_nodes >= 0 && B > 0 && (kind == 0 || kind == 1));
    if (num_nodes == 0) return PJAMD_OK;
    dim3 grid(cdiv(num_nodes, kFwdRows), cdiv(B, 256));
    if (kind == 0)
        input_fwd_kernel<0><<<grid, 256, 0, stream>>>(params, node_mars, data, vids, s_pids, ncats, num_nodes, node_offset, B);
    else
        input_fwd_kernel<1><<<grid, 256, 0, stream>>>(params, node_mars, data, vids, s_pids, ncats, num_nodes, node_offset, B);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

template <typename PF>
static int input_backward_impl(int kind, PF* pflows, const float* node_flows, const int* data, const int* vids,
                               const int64_t* s_pfids, const int* ncats, int64_t num_nodes, int64_t node_offset,
                               int64_t B, bool atomic_out, int kmax, hipStream_t stream) {
    if (kind == 1) {
        bern_bwd_kernel<PF><<<(unsigned)num_nodes, 256, 0, stream>>>(pflows, node_flows, data, vids, s_pfids,
                                                                    node_offset, B, atomic_out);
    } else if (kmax <= 8192) {
        int nt = 8192 / kmax;
        if (nt > 16) nt = 16;
        if (nt < 1) nt = 1;
        size_t lds = (size_t)nt * kmax * sizeof(float);
        cat_bwd_lds_kernel<PF><<<cdiv(num_nodes, nt), 256, lds, stream>>>(pflows, node_flows, data, vids, s_pfids, ncats,
                                                                        num_nodes, node_offset, B, nt, kmax, atomic_out);
    } else {
        dim3 grid((unsigned)num_nodes, cdiv(B, 256));
        cat_bwd_global_kernel<PF><<<grid, 256, 0, stream>>>(pflows, node_flows, data, vids, s_pfids, ncats, num_nodes,
                                                           node_offset, B);
    }
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

// `kmax_or_tied`: bit 0 = some flow ranges are shared by several input vectors (atomic adds);
// bits 1.. = largest category count of the layer (Categorical) .
PJAMD_EXPORT int pjamd_input_backward(int kind, void* pflows, const float* node_flows, int pf64, const int* data,
                                      const int* vids, const int64_t* s_pfids, const int* ncats, int64_t num_nodes,
                                      int64_t node_offset, int64_t B, int kmax_or_tied, hipStream_t stream) {
    PJAMD_REQUIRE(num_nodes >= 0 && B > 0 && (kind == 0 || kind == 1));
    if (num_nodes == 0) return PJAMD_OK;
    const bool atomic_out = (kmax_or_tied & 1) != 0;
    const int kmax = kmax_or_tied >> 1;
    PJAMD_REQUIRE(kind == 1 || kmax > 0);
    if (pf64)
        return input_backward_impl<double>(kind, (double*)pflows, node_flows, data, vids, s_pfids, ncats, num_nodes,
                                           node_offset, B, atomic_out, kmax, stream);
    return input_backward_impl<float>(kind, (float*)pflows, node_flows, data, vids, s_pfids, ncats, num_nodes,
                                      node_offset, B, atomic_out, kmax, stream);
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

PJAMD_EXPORT int pjamd_input_sample(int kind, void* out, const float* params, const uint8_t* node_act,
                                    const float* uniforms, const int* vids, const int64_t* s_pids, const int* ncats,
                                    int out_i64, int64_t num_nodes, int64_t node_offset, int64_t S, hipStream_t stream) {
    PJAMD_REQUIRE(num_nodes >= 0 && S > 0 && (kind == 0 || kind == 1));
    if (num_nodes == 0) return PJAMD_OK;
    dim3 grid((unsigned)num_nodes, cdiv(S, 256));
    if (out_i64)
        input_sample_kernel<int64_t><<<grid, 256, 0, stream>>>(kind, (int64_t*)out, params, node_act, uniforms, vids,
                                                              s_pids, ncats, num_nodes, node_offset, S);
    else
        input_sample_kernel<uint8_t><<<grid, 256, 0, stream>>>(kind, (uint8_t*)out, params, node_act, uniforms, vids,
                                                              s_pids, ncats, num_nodes, node_offset, S);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}
