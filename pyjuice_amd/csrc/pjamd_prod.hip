// Product layers: gather-add over child rows.  Forward: element_mars[nid + j] = sum_k
// node_mars[cid_k + j]; backward (accum = 1): node_flows[u + j] += sum_p element_flows[par_p + j].
// Index 0 marks padding and is skipped.  HBM-bound: rows are read / written as float4 along
// the batch dimension when B is a multiple of 4.
#include "pjamd_common.h"

using namespace pjamd;

namespace {

constexpr int kRows = 4;  // rows (j offsets inside a block) handled by one workgroup

template <int VEC, bool ACCUM>
__global__ __launch_bounds__(256) void gather_add_kernel(float* __restrict__ out, const float* __restrict__ in,
                                                         const int* __restrict__ nids, const int* __restrict__ cids,
                                                         int K, int bs, int64_t B) {
    // blockIdx.x -> (block j, row group); blockIdx.y -> batch chunk of 256 * VEC.
    const int groups = (bs + kRows - 1) / kRows;
    const int j = blockIdx.x / groups;
    const int r0 = (blockIdx.x % groups) * kRows;
    const int64_t b = ((int64_t)blockIdx.y * blockDim.x + threadIdx.x) * VEC;
    if (b >= B) return;
    const int* crow = cids + (int64_t)j * K;
    const int64_t n0 = nids[j];
    using V = typename std::conditional<VEC == 4, float4, float>::type;
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
        const int row = r0 + r;
        if (row >= bs) break;
        float acc[VEC];
#pragma unroll
        for (int q = 0; q < VEC; ++q) acc[q] = 0.0f;
        for (int k = 0; k < K; ++k) {
            const int c = crow[k];
            if (c == 0) continue;
            const float* src = in + ((int64_t)c + row) * B + b;
            if (VEC == 4) {
                const float4 v = *reinterpret_cast<const float4*>(src);
                acc[0] += v.x; acc[1] += v.y; acc[2] += v.z; acc[3] += v.w;
            } else {
                acc[0] += *src;
            }
        }
        float* dst = out + (n0 + row) * B + b;
        if (VEC == 4) {
            float4 o;
            if (ACCUM) {
                o = *reinterpret_cast<float4*>(dst);
                o.x += acc[0]; o.y += acc[1]; o.z += acc[2]; o.w += acc[3];
            } else {
                o = make_float4(acc[0], acc[1], acc[2], acc[3]);
            }
            *reinterpret_cast<float4*>(dst) = o;
        } else {
            if (ACCUM) *dst += acc[0]; else *dst = acc[0];
        }
    }
}

}  // namespace

PJAMD_EXPORT int pjamd_prod_forward(float* out, const float* in, const int* nids, const int* cids, int n_blocks, int K,
                                    int bs, int64_t B, int accum, hipStream_t stream) {
    PJAMD_REQUIRE(n_blocks >= 0 && K >= 1 && bs >= 1 && B > 0);
    if (n_blocks == 0) return PJAMD_OK;
    const int groups = (bs + kRows - 1) / kRows;
    const bool vec = (B % 4 == 0) && (((uintptr_t)out | (uintptr_t)in) % 16 == 0);
    const int64_t lanes = vec ? B / 4 : B;
    dim3 grid((unsigned)((int64_t)n_blocks * groups), cdiv(lanes, 256));
    if (vec) {
        if (accum) gather_add_kernel<4, true><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, bs, B);
        else gather_add_kernel<4, false><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, bs, B);
    } else {
        if (accum) gather_add_kernel<1, true><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, bs, B);
        else gather_add_kernel<1, false><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, bs, B);
    }
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_prod_backward(float* node_flows, const float* element_flows, const int* u_cids,
                                     const int* parids, int n_blocks, int K, int bs, int64_t B, hipStream_t stream) {
    return pjamd_prod_forward(node_flows, element_flows, u_cids, parids, n_blocks, K, bs, B, 1, stream);
}
