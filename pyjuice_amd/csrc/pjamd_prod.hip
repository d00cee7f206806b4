// Product layers: gather-add over child rows.  Forward: element_mars[nid + j] = sum_k
// node_mars[cid_k + j]; backward (accum = 1): node_flows[u + j] += sum_p element_flows[par_p + j].
// Index 0 marks padding and is skipped.
//
// HBM-bound.  One wavefront owns RPW consecutive output rows of one block and 64 float4 (256
// batch columns) of them: the block's K child indices are uniform across the wavefront (scalar
// loads, shared by the RPW rows), and all RPW * K child loads are issued before the first is
// consumed, so a launch costs two dependent memory round trips (indices, then data) whatever K
// is, and a wavefront moves RPW rows' worth of data per round trip.  K is a template parameter
// for the padded widths the compiler produces (powers of two up to 16); wider rows take a
// runtime loop (one row per wavefront).
#include "pjamd_common.h"

using namespace pjamd;

namespace {

constexpr int kWaves = 4;        // wavefronts per workgroup
constexpr int kLoadsInFlight = 16;  // RPW * K float4 child loads per lane

template <int KT, bool VEC, bool ACCUM, int RPW>
__global__ __launch_bounds__(256) void gather_add_kernel(float* __restrict__ out, const float* __restrict__ in,
                                                         const int* __restrict__ nids, const int* __restrict__ cids,
                                                         int K, int lbs, int64_t rows, int64_t B) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row0 = ((int64_t)blockIdx.x * kWaves + w) * RPW;   // RPW divides the block size
    if (row0 >= rows) return;
    const int64_t j = row0 >> lbs;
    const int r0 = (int)(row0 & ((1 << lbs) - 1));
    constexpr int PER = VEC ? 4 : 1;
    const int64_t b = ((int64_t)blockIdx.y * 64 + lane) * PER;
    if (b >= B) return;
    const int* crow = cids + j * K;
    float* dst = out + ((int64_t)nids[j] + r0) * B + b;
    const int kn = KT > 0 ? KT : K;

    if (VEC) {
        if (KT > 0) {
            constexpr int KK = KT > 0 ? KT : 1;
            int c[KK];
#pragma unroll
            for (int k = 0; k < KK; ++k) c[k] = crow[k];
            float4 v[RPW][KK];
#pragma unroll
            for (int q = 0; q < RPW; ++q)
#pragma unroll
                for (int k = 0; k < KK; ++k)
                    v[q][k] = c[k] ? *reinterpret_cast<const float4*>(in + ((int64_t)c[k] + r0 + q) * B + b)
                                   : make_float4(0.f, 0.f, 0.f, 0.f);
            float4 prev[ACCUM ? RPW : 1];
            if (ACCUM) {
#pragma unroll
                for (int q = 0; q < RPW; ++q) prev[q] = *reinterpret_cast<const float4*>(dst + (int64_t)q * B);
            }
#pragma unroll
            for (int q = 0; q < RPW; ++q) {
                float4 acc = ACCUM ? prev[q] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                for (int k = 0; k < KK; ++k) {
                    acc.x += v[q][k].x; acc.y += v[q][k].y; acc.z += v[q][k].z; acc.w += v[q][k].w;
                }
                *reinterpret_cast<float4*>(dst + (int64_t)q * B) = acc;
            }
        } else {
            float4 acc = ACCUM ? *reinterpret_cast<const float4*>(dst) : make_float4(0.f, 0.f, 0.f, 0.f);
            for (int k = 0; k < kn; ++k) {
                const int c = crow[k];
                if (c == 0) continue;
                const float4 v = *reinterpret_cast<const float4*>(in + ((int64_t)c + r0) * B + b);
                acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
            }
            *reinterpret_cast<float4*>(dst) = acc;
        }
    } else {
        float acc = ACCUM ? *dst : 0.0f;
        if (KT > 0) {
            float v[KT > 0 ? KT : 1];
#pragma unroll
            for (int k = 0; k < (KT > 0 ? KT : 1); ++k) {
                const int c = crow[k];
                v[k] = c ? in[((int64_t)c + r0) * B + b] : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < (KT > 0 ? KT : 1); ++k) acc += v[k];
        } else {
            for (int k = 0; k < kn; ++k) {
                const int c = crow[k];
                if (c != 0) acc += in[((int64_t)c + r0) * B + b];
            }
        }
        *dst = acc;
    }
}

// Rows per wavefront for K children: kLoadsInFlight / K float4 loads per row group, at most the
// block size (16-byte form only; 1 otherwise), and only for launches of at least kManyWaves
// one-row wavefronts: a small launch is a few dependent round trips, and spreading its loads
// over more wavefronts (one row each) finishes sooner.
constexpr int64_t kManyWaves = 32768;

template <int KT, bool VEC, bool ACCUM>
void launch_k(float* out, const float* in, const int* nids, const int* cids, int K, int lbs, int64_t rows, int64_t B,
              unsigned gy, hipStream_t stream) {
    constexpr int R = (VEC && KT > 0) ? (kLoadsInFlight / KT > 1 ? kLoadsInFlight / KT : 1) : 1;
    int rpw = rows * (int64_t)gy >= kManyWaves ? R : 1;
    while (rpw > 1 && rpw > (1 << lbs)) rpw >>= 1;
    const int64_t waves = cdiv(rows, (int64_t)rpw);
    const dim3 grid((unsigned)cdiv(waves, kWaves), gy);
    switch (rpw) {
        case 16: gather_add_kernel<KT, VEC, ACCUM, (R >= 16 ? 16 : 1)><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, lbs, rows, B); break;
        case 8: gather_add_kernel<KT, VEC, ACCUM, (R >= 8 ? 8 : 1)><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, lbs, rows, B); break;
        case 4: gather_add_kernel<KT, VEC, ACCUM, (R >= 4 ? 4 : 1)><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, lbs, rows, B); break;
        case 2: gather_add_kernel<KT, VEC, ACCUM, (R >= 2 ? 2 : 1)><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, lbs, rows, B); break;
        default: gather_add_kernel<KT, VEC, ACCUM, 1><<<grid, 256, 0, stream>>>(out, in, nids, cids, K, lbs, rows, B); break;
    }
}

template <bool VEC, bool ACCUM>
void launch_gather(float* out, const float* in, const int* nids, const int* cids, int K, int lbs, int64_t rows,
                   int64_t B, unsigned gy, hipStream_t stream) {
    switch (K) {
        case 1: launch_k<1, VEC, ACCUM>(out, in, nids, cids, K, lbs, rows, B, gy, stream); break;
        case 2: launch_k<2, VEC, ACCUM>(out, in, nids, cids, K, lbs, rows, B, gy, stream); break;
        case 4: launch_k<4, VEC, ACCUM>(out, in, nids, cids, K, lbs, rows, B, gy, stream); break;
        case 8: launch_k<8, VEC, ACCUM>(out, in, nids, cids, K, lbs, rows, B, gy, stream); break;
        case 16: launch_k<16, VEC, ACCUM>(out, in, nids, cids, K, lbs, rows, B, gy, stream); break;
        default: launch_k<0, VEC, ACCUM>(out, in, nids, cids, K, lbs, rows, B, gy, stream); break;
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
    PJAMD_REQUIRE(cdiv(rows, kWaves) < (int64_t)1 << 31);
    const unsigned gy = (unsigned)cdiv(B, per_wave);
    if (vec) {
        if (accum) launch_gather<true, true>(out, in, nids, cids, K, lbs, rows, B, gy, stream);
        else launch_gather<true, false>(out, in, nids, cids, K, lbs, rows, B, gy, stream);
    } else {
        if (accum) launch_gather<false, true>(out, in, nids, cids, K, lbs, rows, B, gy, stream);
        else launch_gather<false, false>(out, in, nids, cids, K, lbs, rows, B, gy, stream);
    }
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_prod_backward(float* node_flows, const float* element_flows, const int* u_cids,
                                     const int* parids, int n_blocks, int K, int bs, int64_t B, hipStream_t stream) {
    return pjamd_prod_forward(node_flows, element_flows, u_cids, parids, n_blocks, K, bs, B, 1, stream);
}
