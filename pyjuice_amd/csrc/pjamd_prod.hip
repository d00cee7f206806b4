// Product layers: gather-add over child rows.  Forward: element_mars[nid + j] = sum_k
// node_mars[cid_k + j]; backward (accum = 1): node_flows[u + j] += sum_p element_flows[par_p + j].
// Index 0 marks padding and is skipped.
//
// HBM-bound.  One wavefront owns one output row and 64 float4 (256 batch columns) of it: the
// row's K child indices are uniform across the wavefront (scalar loads), and all K child
// loads are issued before the first is consumed, so a launch costs two dependent memory round
// trips (indices, then data) whatever K is.  K is a template parameter for the padded widths
// the compiler produces (powers of two up to 16); wider rows take a runtime loop.
#include "pjamd_common.h"

using namespace pjamd;

namespace {

constexpr int kWaves = 4;  // rows per workgroup (one per wavefront)

template <int KT, bool VEC, bool ACCUM>
__global__ __launch_bounds__(256) void gather_add_kernel(float* __restrict__ out, const float* __restrict__ in,
                                                         const int* __restrict__ nids, const int* __restrict__ cids,
                                                         int K, int lbs, int64_t rows, int64_t B) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * kWaves + w;
    if (row >= rows) return;
    const int64_t j = row >> lbs;
    const int r = (int)(row & ((1 << lbs) - 1));
    constexpr int PER = VEC ? 4 : 1;
    const int64_t b = ((int64_t)blockIdx.y * 64 + lane) * PER;
    if (b >= B) return;
    const int* crow = cids + j * K;
    float* dst = out + ((int64_t)nids[j] + r) * B + b;
    const int kn = KT > 0 ? KT : K;

    if (VEC) {
        float4 acc = ACCUM ? *reinterpret_cast<const float4*>(dst) : make_float4(0.f, 0.f, 0.f, 0.f);
        if (KT > 0) {
            float4 v[KT > 0 ? KT : 1];
#pragma unroll
            for (int k = 0; k < (KT > 0 ? KT : 1); ++k) {
                const int c = crow[k];
                v[k] = c ? *reinterpret_cast<const float4*>(in + ((int64_t)c + r) * B + b) : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int k = 0; k < (KT > 0 ? KT : 1); ++k) {
                acc.x += v[k].x; acc.y += v[k].y; acc.z += v[k].z; acc.w += v[k].w;
            }
        } else {
            for (int k = 0; k < kn; ++k) {
                const int c = crow[k];
                if (c == 0) continue;
                const float4 v = *reinterpret_cast<const float4*>(in + ((int64_t)c + r) * B + b);
                acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
            }
        }
        *reinterpret_cast<float4*>(dst) = acc;
    } else {
        float acc = ACCUM ? *dst : 0.0f;
        if (KT > 0) {
            float v[KT > 0 ? KT : 1];
#pragma unroll
            for (int k = 0; k < (KT > 0 ? KT : 1); ++k) {
                const int c = crow[k];
                v[k] = c ? in[((int64_t)c + r) * B + b] : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < (KT > 0 ? KT : 1); ++k) acc += v[k];
        } else {
            for (int k = 0; k < kn; ++k) {
                const int c = crow[k];
                if (c != 0) acc += in[((int64_t)c + r) * B + b];
            }
        }
        *dst = acc;
    }
}

template <bool VEC, bool ACCUM>
void launch_gather(float* out, const float* in, const int* nids, const int* cids, int K, int lbs, int64_t rows,
                   int64_t B, dim3 grid, hipStream_t stream) {
    switch (K) {
        case 1: gather_add_kernel<1, VEC, ACCUM><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, lbs, rows, B); break;
        case 2: gather_add_kernel<2, VEC, ACCUM><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, lbs, rows, B); break;
        case 4: gather_add_kernel<4, VEC, ACCUM><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, lbs, rows, B); break;
        case 8: gather_add_kernel<8, VEC, ACCUM><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, lbs, rows, B); break;
        case 16: gather_add_kernel<16, VEC, ACCUM><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, lbs, rows, B); break;
        default: gather_add_kernel<0, VEC, ACCUM><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, lbs, rows, B); break;
    }
}

int log2_exact(int v) {
    if (v < 1 || (v & (v - 1)) != 0) return -1;
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

}  // namespace

PJAMD_EXPORT int pjamd_prod_forward(float* out, const float* in, const int* nids, const int* cids, int n_blocks, int K,
                                    int bs, int64_t B, int accum, hipStream_t stream) {
    PJAMD_REQUIRE(n_blocks >= 0 && K >= 1 && bs >= 1 && B > 0);
    const int lbs = log2_exact(bs);
    PJAMD_REQUIRE(lbs >= 0);
    if (n_blocks == 0) return PJAMD_OK;
    const int64_t rows = (int64_t)n_blocks * bs;
    const bool vec = (B % 4 == 0) && (((uintptr_t)out | (uintptr_t)in) % 16 == 0);
    const int64_t per_wave = vec ? 256 : 64;
    const int64_t gx = cdiv(rows, kWaves);
    PJAMD_REQUIRE(gx < (int64_t)1 << 31);
    dim3 grid((unsigned)gx, (unsigned)cdiv(B, per_wave));
    if (vec) {
        if (accum) launch_gather<true, true>(out, in, nids, cids, K, lbs, rows, B, grid, stream);
        else launch_gather<true, false>(out, in, nids, cids, K, lbs, rows, B, grid, stream);
    } else {
        if (accum) launch_gather<false, true>(out, in, nids, cids, K, lbs, rows, B, grid, stream);
        else launch_gather<false, false>(out, in, nids, cids, K, lbs, rows, B, grid, stream);
    }
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_prod_backward(float* node_flows, const float* element_flows, const int* u_cids,
                                     const int* parids, int n_blocks, int K, int bs, int64_t B, hipStream_t stream) {
    return pjamd_prod_forward(node_flows, element_flows, u_cids, parids, n_blocks, K, bs, B, 1, stream);
}
