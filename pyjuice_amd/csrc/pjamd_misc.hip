// Buffer fills, data transposition, sum-parameter EM, tied-flow fusing, top-down sampling
// of sum / product layers, and the library's error strings.
#include "pjamd_common.h"

using namespace pjamd;

namespace {

__global__ __launch_bounds__(256) void fill_kernel(float* __restrict__ p, float v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// p[i] = log(p[i]) (-inf for 0), grid-stride.
__global__ __launch_bounds__(256) void log_kernel(float* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = logf(p[i]);
}

// [B, V] integer data of 1, 4 or 8 bytes -> [V, B] int32 through a 32 x 32 LDS tile.
template <typename T>
__global__ __launch_bounds__(256) void to_vb_kernel(int* __restrict__ out, const T* __restrict__ src, int64_t B, int64_t V) {
    __shared__ int tile[32][33];
    const int64_t v0 = (int64_t)blockIdx.x * 32, b0 = (int64_t)blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int64_t b = b0 + r, v = v0 + tx;
        tile[r][tx] = (b < B && v < V) ? (int)src[b * V + v] : 0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t v = v0 + r, b = b0 + tx;
        if (b < B && v < V) out[v * B + b] = tile[tx][r];
    }
}

// Mini-batch EM of the sum parameters.  Record = [par_start, pflow_start, node_start,
// block_size, n_children]; the parameter of child c of node i sits at par_start + c * block_size
// + i (same for flows).  A workgroup of kEmWaves wavefronts handles (record, up to 64 nodes).
// Scalar path (any tile): lanes run along the nodes (ti nodes, G = 64 / ti lanes per node) and
// the wavefronts split the children.  When every lane owns at most kRegChildren children (a tile
// of nodes dividing 64, and at most kRegChildren children per lane: 256 children at block size
// >= 64, 512 at 32) the flows and parameters stay in registers between the normaliser and the
// update, so each is read from memory once; otherwise the update re-reads them.  The 16-byte
// path below takes the tiles whose rows allow it.
constexpr int kEmWaves = 16;
constexpr int kRegChildren = 16;
constexpr int kEmChunk = 8;       // children per lane per round trip when they do not fit registers

template <typename PF>
__device__ __forceinline__ void sum_em_scalar(float* __restrict__ params, const PF* __restrict__ pflows, int64_t par0,
                                              int64_t pf0, int bs, int nch, int tn, double step, double pc,
                                              bool keep_zero, double (&part)[kEmWaves][64]) {
    const int ti = tn;                      // nodes of this workgroup: min(64, bs - 64 blockIdx.y)
    const int G = 64 / ti;                  // lanes per node inside a wavefront
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.y * 64 + (lane % ti);
    const int g = lane / ti + G * w;        // this lane's child offset
    const int stride = G * kEmWaves;
    const bool lane_ok = (lane / ti) < G;   // lanes past G * ti (when 64 % ti != 0) idle
    const double npc = pc / (double)nch;
    const bool in_regs = (64 % ti == 0) && (nch <= stride * kRegChildren);

    double fr[kRegChildren];
    float tr[kRegChildren];
    double cum = 0.0;
    if (in_regs) {
#pragma unroll
        for (int q = 0; q < kRegChildren; ++q) {
            const int c = g + q * stride;
            const int64_t off = (int64_t)c * bs + i;
            fr[q] = (c < nch) ? (double)pflows[pf0 + off] : 0.0;
            tr[q] = (c < nch) ? params[par0 + off] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < kRegChildren; ++q) {
            const int c = g + q * stride;
            if (c < nch) cum += keep_zero ? ((tr[q] < 1e-12f) ? 0.0 : fr[q] + npc) : fr[q];
        }
    } else if (lane_ok) {
        // kEmChunk children per lane per round trip (clamped to a real child, masked on use)
        for (int c0 = g; c0 < nch; c0 += kEmChunk * stride) {
            double f[kEmChunk];
            float t[kEmChunk];
#pragma unroll
            for (int q = 0; q < kEmChunk; ++q) {
                const int c = c0 + q * stride;
                const int64_t off = (int64_t)(c < nch ? c : c0) * bs + i;
                f[q] = (double)pflows[pf0 + off];
                t[q] = params[par0 + off];
            }
#pragma unroll
            for (int q = 0; q < kEmChunk; ++q)
                if (c0 + q * stride < nch) cum += keep_zero ? ((t[q] < 1e-12f) ? 0.0 : f[q] + npc) : f[q];
        }
    }
    // reduce over the G lane groups of a wavefront, then over the wavefronts
    for (int o = ti; o < 64; o <<= 1) cum += __shfl_xor(cum, o, 64);
    part[w][lane] = cum;
    __syncthreads();
    double tot = 0.0;
#pragma unroll
    for (int q = 0; q < kEmWaves; ++q) tot += part[q][lane % ti];
    const double denom = keep_zero ? tot : tot + pc;
    if (in_regs) {
#pragma unroll
        for (int q = 0; q < kRegChildren; ++q) {
            const int c = g + q * stride;
            if (c < nch) {
                const double th = (double)tr[q];
                double upd = (1.0 - step) * th + step * ((fr[q] + npc) / denom);
                if (keep_zero && th < 1e-12) upd = 0.0;
                params[par0 + (int64_t)c * bs + i] = (float)upd;
            }
        }
    } else if (lane_ok) {
        for (int c0 = g; c0 < nch; c0 += kEmChunk * stride) {
            double f[kEmChunk];
            float t[kEmChunk];
#pragma unroll
            for (int q = 0; q < kEmChunk; ++q) {
                const int c = c0 + q * stride;
                const int64_t off = (int64_t)(c < nch ? c : c0) * bs + i;
                f[q] = (double)pflows[pf0 + off];
                t[q] = params[par0 + off];
            }
#pragma unroll
            for (int q = 0; q < kEmChunk; ++q) {
                const int c = c0 + q * stride;
                if (c < nch) {
                    const double th = (double)t[q];
                    double upd = (1.0 - step) * th + step * ((f[q] + npc) / denom);
                    if (keep_zero && th < 1e-12) upd = 0.0;
                    params[par0 + (int64_t)c * bs + i] = (float)upd;
                }
            }
        }
    }
}

// The same update with 16-byte accesses, for a tile of tn = min(64, bs - i0) nodes where tn is a
// power of two and every row start is 16-byte aligned: a lane owns 4 consecutive nodes (quad q)
// of kVecChildren children (lanes of one quad in a wavefront, then the wavefronts, split the
// children), loading a float4 of parameters and 16 bytes of flows per child and row; the
// normaliser is reduced over the lanes of a quad (shuffles), then over the wavefronts (LDS).
// A quarter of the scalar path's memory instructions for the same bytes.
constexpr int kVecChildren = 4;

template <typename PF>
struct Flow4 {   // four consecutive flows of one child, loaded as 16-byte vectors
    double v[4];
    __device__ __forceinline__ void load(const PF* p) {
        if constexpr (sizeof(PF) == 8) {
            const double2 a = *reinterpret_cast<const double2*>(p), b = *reinterpret_cast<const double2*>(p + 2);
            v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
        } else {
            const float4 a = *reinterpret_cast<const float4*>(p);
            v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        }
    }
};

template <typename PF>
__device__ __forceinline__ void sum_em_vec(float* __restrict__ params, const PF* __restrict__ pflows, int64_t par0,
                                           int64_t pf0, int bs, int nch, int tn, double step, double pc,
                                           bool keep_zero, double (&part)[kEmWaves][64], double* den) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i0 = blockIdx.y * 64;
    const int QT = tn / 4;                 // node quads of the tile (1 .. 16, a power of two)
    const int GW = 64 / QT;                // lanes per quad in a wavefront
    const int q = lane % QT;
    const int g = lane / QT + GW * w;      // this lane's first child
    const int stride = GW * kEmWaves;
    const bool in_regs = nch <= stride * kVecChildren;
    const double npc = pc / (double)nch;
    const int64_t col = i0 + 4 * q;
    auto contrib = [&](double f, float t) { return keep_zero ? ((t < 1e-12f) ? 0.0 : f + npc) : f; };

    float4 tr[kVecChildren];
    Flow4<PF> fr[kVecChildren];
    double cum[4] = {0.0, 0.0, 0.0, 0.0};
    if (in_regs) {
#pragma unroll
        for (int u = 0; u < kVecChildren; ++u) {
            const int c = g + u * stride;
            const int cc = c < nch ? c : g < nch ? g : 0;   // clamped to a real child, masked below
            const int64_t off = (int64_t)cc * bs + col;
            tr[u] = *reinterpret_cast<const float4*>(params + par0 + off);
            fr[u].load(pflows + pf0 + off);
        }
#pragma unroll
        for (int u = 0; u < kVecChildren; ++u) {
            if (g + u * stride >= nch) continue;
            cum[0] += contrib(fr[u].v[0], tr[u].x);
            cum[1] += contrib(fr[u].v[1], tr[u].y);
            cum[2] += contrib(fr[u].v[2], tr[u].z);
            cum[3] += contrib(fr[u].v[3], tr[u].w);
        }
    } else {
        for (int c0 = g; c0 < nch; c0 += kVecChildren * stride) {
            float4 t[kVecChildren];
            Flow4<PF> f[kVecChildren];
#pragma unroll
            for (int u = 0; u < kVecChildren; ++u) {
                const int c = c0 + u * stride;
                const int64_t off = (int64_t)(c < nch ? c : c0) * bs + col;
                t[u] = *reinterpret_cast<const float4*>(params + par0 + off);
                f[u].load(pflows + pf0 + off);
            }
#pragma unroll
            for (int u = 0; u < kVecChildren; ++u) {
                if (c0 + u * stride >= nch) continue;
                cum[0] += contrib(f[u].v[0], t[u].x);
                cum[1] += contrib(f[u].v[1], t[u].y);
                cum[2] += contrib(f[u].v[2], t[u].z);
                cum[3] += contrib(f[u].v[3], t[u].w);
            }
        }
    }
    // over the lanes of a quad (lane bits >= log2 QT), then over the wavefronts
#pragma unroll
    for (int i = 0; i < 4; ++i)
        for (int o = QT; o < 64; o <<= 1) cum[i] += __shfl_xor(cum[i], o, 64);
    if (lane < QT) {
#pragma unroll
        for (int i = 0; i < 4; ++i) part[w][4 * q + i] = cum[i];
    }
    __syncthreads();
    if (threadIdx.x < tn) {
        double tot = 0.0;
#pragma unroll
        for (int v = 0; v < kEmWaves; ++v) tot += part[v][threadIdx.x];
        den[threadIdx.x] = keep_zero ? tot : tot + pc;
    }
    __syncthreads();
    double dn[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) dn[i] = den[4 * q + i];
    auto upd = [&](double f, float t, double d) {
        const double th = (double)t;
        double nv = (1.0 - step) * th + step * ((f + npc) / d);
        if (keep_zero && th < 1e-12) nv = 0.0;
        return (float)nv;
    };
    if (in_regs) {
#pragma unroll
        for (int u = 0; u < kVecChildren; ++u) {
            const int c = g + u * stride;
            if (c >= nch) continue;
            const float4 o = make_float4(upd(fr[u].v[0], tr[u].x, dn[0]), upd(fr[u].v[1], tr[u].y, dn[1]),
                                         upd(fr[u].v[2], tr[u].z, dn[2]), upd(fr[u].v[3], tr[u].w, dn[3]));
            *reinterpret_cast<float4*>(params + par0 + (int64_t)c * bs + col) = o;
        }
    } else {
        for (int c0 = g; c0 < nch; c0 += kVecChildren * stride) {
            float4 t[kVecChildren];
            Flow4<PF> f[kVecChildren];
#pragma unroll
            for (int u = 0; u < kVecChildren; ++u) {
                const int c = c0 + u * stride;
                const int64_t off = (int64_t)(c < nch ? c : c0) * bs + col;
                t[u] = *reinterpret_cast<const float4*>(params + par0 + off);
                f[u].load(pflows + pf0 + off);
            }
#pragma unroll
            for (int u = 0; u < kVecChildren; ++u) {
                const int c = c0 + u * stride;
                if (c >= nch) continue;
                const float4 o = make_float4(upd(f[u].v[0], t[u].x, dn[0]), upd(f[u].v[1], t[u].y, dn[1]),
                                             upd(f[u].v[2], t[u].z, dn[2]), upd(f[u].v[3], t[u].w, dn[3]));
                *reinterpret_cast<float4*>(params + par0 + (int64_t)c * bs + col) = o;
            }
        }
    }
}

// One workgroup per (record, tile of 64 nodes): the 16-byte path where the tile allows it
// (vec_ptrs: both base pointers 16-byte aligned), the scalar path otherwise.
template <typename PF>
__global__ __launch_bounds__(64 * kEmWaves) void sum_em_kernel(float* __restrict__ params, const PF* __restrict__ pflows,
                                                     const int* __restrict__ recs, int n_recs, double step, double pc,
                                                     bool keep_zero, bool vec_ptrs) {
    __shared__ double part[kEmWaves][64];
    __shared__ double den[64];
    const int rec = blockIdx.x;
    const int* R = recs + (int64_t)rec * 5;
    const int64_t par0 = R[0], pf0 = R[1];
    const int bs = R[3], nch = R[4];
    if (blockIdx.y * 64 >= bs || nch <= 0) return;   // uniform over the workgroup
    const int tn = bs - (int)blockIdx.y * 64 < 64 ? bs - (int)blockIdx.y * 64 : 64;
    constexpr int kPfAlign = 16 / (int)sizeof(PF);
    const bool vec = vec_ptrs && bs % 4 == 0 && tn >= 4 && (tn & (tn - 1)) == 0 && par0 % 4 == 0 &&
                     pf0 % kPfAlign == 0;
    if (vec)
        sum_em_vec<PF>(params, pflows, par0, pf0, bs, nch, tn, step, pc, keep_zero, part, den);
    else
        sum_em_scalar<PF>(params, pflows, par0, pf0, bs, nch, tn, step, pc, keep_zero, part);
}

// pf[target + j] = sum over copies c of pf[copy_c + j] (the copies include the target range).
template <typename PF>
__global__ __launch_bounds__(256) void fuse_kernel(PF* __restrict__ pf, int64_t target, int64_t n,
                                                   const int64_t* __restrict__ copies, int ncopies) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    PF s = 0;
    for (int c = 0; c < ncopies; ++c) s += pf[copies[c] + j];
    pf[target + j] = s;
}

// Sum layer top-down step.  A workgroup takes one node and 256 samples; if any of them is active
// (its flag is cleared as it is read), the workgroup stages the node's child parameters in LDS in
// chunks of kSampleChunk (the reference's BLOCK_K of at most 512, sample.py:115-131), one thread
// turns a chunk into its FP32 running sums (left to right, from 0), and every active sample
// binary-searches them: the pick is the first child whose running sum exceeds u; a u at or past
// the chunk's total moves on to the next chunk with the total subtracted; past the last chunk the
// last child is taken.  The running sums are the sequential ones, so the picks equal those of a
// per-sample left-to-right scan.
constexpr int kSampleChunk = 512;

__device__ __forceinline__ void serial_prefix(float* P, int m) {
    float c = 0.0f;
    int k = 0;
    for (; k + 8 <= m; k += 8) {
        float w[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) w[q] = P[k + q];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            c += w[q];
            P[k + q] = c;
        }
    }
    for (; k < m; ++k) {
        c += P[k];
        P[k] = c;
    }
}

__global__ __launch_bounds__(256) void sample_sum_kernel(uint8_t* __restrict__ elem_act, uint8_t* __restrict__ node_act,
                                                         const float* __restrict__ uniforms,
                                                         const uint32_t* __restrict__ seed,
                                                         const float* __restrict__ params, const float* __restrict__ cdf,
                                                         const int* __restrict__ nids,
                                                         const int* __restrict__ cids, const int* __restrict__ pids,
                                                         const int* __restrict__ nch, int K, int bs, int64_t S) {
    __shared__ float P[kSampleChunk];
    const int tid = threadIdx.x;
    const int64_t s = (int64_t)blockIdx.y * blockDim.x + tid;
    const int j = blockIdx.x / bs, i = blockIdx.x % bs;
    const int64_t node = (int64_t)nids[j] + i;
    const int64_t o = node * S + s;
    const bool act = s < S && node_act[o] != 0;
    if (!__syncthreads_or(act)) return;
    float u = 0.0f;
    if (act) {
        node_act[o] = 0;
        u = uniforms != nullptr ? uniforms[o] : philox_uniform(seed[0], seed[1], (uint32_t)s, (uint32_t)node);
    }
    const int n = nch[j];
    const int* crow = cids + (int64_t)j * K;
    const int* prow = pids + (int64_t)j * K;
    int pick = -1;
    for (int c0 = 0; c0 < n; c0 += kSampleChunk) {
        const int m = min(kSampleChunk, n - c0);
        if (cdf != nullptr) {   // running sums precomputed (pjamd_sample_prefix)
            for (int k = tid; k < m; k += blockDim.x) P[k] = cdf[prow[c0 + k] + i];
            __syncthreads();
        } else {
            for (int k = tid; k < m; k += blockDim.x) P[k] = params[prow[c0 + k] + i];
            __syncthreads();
            if (tid == 0) serial_prefix(P, m);
            __syncthreads();
        }
        if (act && pick < 0) {
            const float tot = P[m - 1];
            if (u < tot) {
                int lo = 0, hi = m - 1;        // first k with u < P[k]; P[hi] > u
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (u < P[mid]) hi = mid;
                    else lo = mid + 1;
                }
                pick = c0 + lo;
            } else {
                u -= tot;
            }
        }
        if (!__syncthreads_or(act && pick < 0)) break;   // also orders the reads before the next chunk
    }
    if (act) elem_act[(int64_t)crow[pick < 0 ? n - 1 : pick] * S + s] = 1;
}

// Running sums of rows r < n_rows: out[base_r + c * stride_r] = sum_{c' <= c} in[base_r + c' *
// stride_r] in FP32, left to right, restarting at every multiple of `chunk` (0: never).  One thread
// per row, kSampleBatch loads in flight.
constexpr int kPrefixBatch = 16;

__global__ __launch_bounds__(256) void row_prefix_kernel(float* __restrict__ out, const float* __restrict__ in,
                                                         const int64_t* __restrict__ base, const int* __restrict__ stride,
                                                         const int* __restrict__ len, int64_t n_rows, int chunk) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    const int64_t b0 = base[r];
    const int st = stride[r], n = len[r];
    float c = 0.0f;
    for (int k0 = 0; k0 < n; k0 += kPrefixBatch) {
        float w[kPrefixBatch];
#pragma unroll
        for (int q = 0; q < kPrefixBatch; ++q) w[q] = in[b0 + (int64_t)min(k0 + q, n - 1) * st];
#pragma unroll
        for (int q = 0; q < kPrefixBatch; ++q) {
            const int k = k0 + q;
            if (k < n) {
                if (chunk > 0 && k % chunk == 0) c = 0.0f;
                c += w[q];
                out[b0 + (int64_t)k * st] = c;
            }
        }
    }
}

// Product layer top-down step: an active element activates all its children (and is cleared).
__global__ __launch_bounds__(256) void sample_prod_kernel(uint8_t* __restrict__ node_act, uint8_t* __restrict__ elem_act,
                                                          const int* __restrict__ nids, const int* __restrict__ cids,
                                                          int K, int bs, int64_t S) {
    const int64_t s = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const int j = blockIdx.x / bs, i = blockIdx.x % bs;
    const int64_t e = ((int64_t)nids[j] + i) * S + s;
    if (!elem_act[e]) return;
    elem_act[e] = 0;
    const int* crow = cids + (int64_t)j * K;
    for (int k = 0; k < K; ++k) {
        const int c = crow[k];
        if (c != 0) node_act[((int64_t)c + i) * S + s] = 1;
    }
}

}  // namespace

PJAMD_EXPORT int pjamd_abi_version() { return 1; }

PJAMD_EXPORT const char* pjamd_error_string(int code) {
    if (code == PJAMD_OK) return "success";
    if (code == PJAMD_EINVAL) return "invalid argument (size, block size or unsupported combination)";
    if (code == PJAMD_EUNSUPPORTED) return "unsupported configuration";
    if (code == PJAMD_ETOOLARGE)
        return "request exceeds the fixed device workspace of the allocation-free entry point "
               "(PJAMD_SPLIT_SCRATCH_FLOATS floats)";
    if (code > 0) return hipGetErrorString((hipError_t)code);
    return "unknown error";
}

PJAMD_EXPORT int pjamd_flows_to_log(float* p, int64_t n, hipStream_t stream) {
    PJAMD_REQUIRE(n >= 0);
    if (n == 0) return PJAMD_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    log_kernel<<<(unsigned)blocks, 256, 0, stream>>>(p, n);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_fill_f32(float* p, float v, int64_t n, hipStream_t stream) {
    PJAMD_REQUIRE(n >= 0);
    if (n == 0) return PJAMD_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    fill_kernel<<<(unsigned)blocks, 256, 0, stream>>>(p, v, n);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_data_to_vb(int* out, int src_code, const void* src, int64_t B, int64_t V, hipStream_t stream) {
    PJAMD_REQUIRE(B > 0 && V > 0);
    dim3 grid(cdiv(V, 32), cdiv(B, 32));
    switch (src_code) {
        case 0: to_vb_kernel<uint8_t><<<grid, 256, 0, stream>>>(out, (const uint8_t*)src, B, V); break;
        case 1: to_vb_kernel<int32_t><<<grid, 256, 0, stream>>>(out, (const int32_t*)src, B, V); break;
        case 2: to_vb_kernel<int64_t><<<grid, 256, 0, stream>>>(out, (const int64_t*)src, B, V); break;
        case 3: to_vb_kernel<int8_t><<<grid, 256, 0, stream>>>(out, (const int8_t*)src, B, V); break;
        default: return PJAMD_EINVAL;
    }
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

// data[v][b] <- -1 (marginalized) where the missing mask is set.  mode 0: mask[v]; 1: mask[b][v];
// 2: mask[v][b].
__global__ __launch_bounds__(256) void data_mask_kernel(int* __restrict__ data, const uint8_t* __restrict__ mask,
                                                        int mode, int64_t B, int64_t V) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t v = blockIdx.y;
    if (b >= B) return;
    const uint8_t m = mode == 0 ? mask[v] : (mode == 1 ? mask[b * V + v] : mask[v * B + b]);
    if (m) data[v * B + b] = -1;
}

PJAMD_EXPORT int pjamd_data_mask(int* data, const uint8_t* mask, int mode, int64_t B, int64_t V, hipStream_t stream) {
    PJAMD_REQUIRE(B > 0 && V > 0 && V < 65536 && mode >= 0 && mode <= 2);
    data_mask_kernel<<<dim3(cdiv(B, 256), (unsigned)V), 256, 0, stream>>>(data, mask, mode, B, V);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_sum_em(float* params, const void* pflows, int pf64, const int* recs, int n_recs, float step,
                              float pseudocount, int keep_zero, int max_bs, hipStream_t stream) {
    PJAMD_REQUIRE(n_recs >= 0 && max_bs >= 1);
    if (n_recs == 0) return PJAMD_OK;
    dim3 grid((unsigned)n_recs, cdiv(max_bs, 64));
    const bool vec_ptrs = ((((uintptr_t)params) | ((uintptr_t)pflows)) % 16) == 0;
    if (pf64)
        sum_em_kernel<double><<<grid, 64 * kEmWaves, 0, stream>>>(params, (const double*)pflows, recs, n_recs, step,
                                                                  pseudocount, keep_zero != 0, vec_ptrs);
    else
        sum_em_kernel<float><<<grid, 64 * kEmWaves, 0, stream>>>(params, (const float*)pflows, recs, n_recs, step,
                                                                 pseudocount, keep_zero != 0, vec_ptrs);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_pflow_fuse(void* pflows, int pf64, int64_t target, int64_t n, const int64_t* copies, int ncopies,
                                  hipStream_t stream) {
    PJAMD_REQUIRE(n >= 0 && ncopies >= 1);
    if (n == 0) return PJAMD_OK;
    if (pf64) fuse_kernel<double><<<cdiv(n, 256), 256, 0, stream>>>((double*)pflows, target, n, copies, ncopies);
    else fuse_kernel<float><<<cdiv(n, 256), 256, 0, stream>>>((float*)pflows, target, n, copies, ncopies);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_sample_sum(uint8_t* elem_act, uint8_t* node_act, const float* uniforms, const uint32_t* seed,
                                  const float* params, const float* cdf, const int* nids, const int* cids,
                                  const int* pids, const int* nch, int n_blocks, int K, int bs, int64_t S,
                                  hipStream_t stream) {
    PJAMD_REQUIRE(n_blocks >= 0 && K >= 1 && bs >= 1 && S > 0 && (uniforms != nullptr || seed != nullptr));
    if (n_blocks == 0) return PJAMD_OK;
    dim3 grid((unsigned)((int64_t)n_blocks * bs), cdiv(S, 256));
    sample_sum_kernel<<<grid, 256, 0, stream>>>(elem_act, node_act, uniforms, seed, params, cdf, nids, cids, pids, nch,
                                                K, bs, S);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_sample_prefix(float* out, const float* in, const int64_t* row_base, const int* row_stride,
                                     const int* row_len, int64_t n_rows, int chunk, hipStream_t stream) {
    PJAMD_REQUIRE(n_rows >= 0 && chunk >= 0);
    if (n_rows == 0) return PJAMD_OK;
    PJAMD_REQUIRE(cdiv(n_rows, 256) < (1u << 31));
    row_prefix_kernel<<<cdiv(n_rows, 256), 256, 0, stream>>>(out, in, row_base, row_stride, row_len, n_rows, chunk);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_sample_prod(uint8_t* node_act, uint8_t* elem_act, const int* nids, const int* cids,
                                   int n_blocks, int K, int bs, int64_t S, hipStream_t stream) {
    PJAMD_REQUIRE(n_blocks >= 0 && K >= 1 && bs >= 1 && S > 0);
    if (n_blocks == 0) return PJAMD_OK;
    dim3 grid((unsigned)((int64_t)n_blocks * bs), cdiv(S, 256));
    sample_prod_kernel<<<grid, 256, 0, stream>>>(node_act, elem_act, nids, cids, K, bs, S);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}
