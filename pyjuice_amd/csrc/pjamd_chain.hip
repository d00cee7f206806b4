// Chains of narrow layers in one launch.
//
// Deep tree-shaped circuits (HCLT, HMM) end in long runs of layers with only a few node blocks:
// one launch per layer keeps most of the GPU idle and pays a launch plus several dependent memory
// round trips per layer.  Every batch column of a circuit pass is independent of every other, so
// a run of layers can be split over the batch instead: workgroup g owns columns
// [32 g, 32 g + 32) and walks the run's operations in order, all node rows of each layer, with a
// workgroup barrier between operations.  No workgroup ever waits for another.
//
// Forward operations (pjamd_chain_forward):
//   kind 0  product layer partition: element_mars[slab + nid + i] = sum_k node_mars[cid_k + i]
//   kind 1  sum layer partition:     node_mars[nid + i] = log sum_k theta[pid_k + i] exp(e_k)
// Backward operations (pjamd_chain_backward):
//   kind 2  sum layer backward partition fused with the product layer below it (children with a
//           single parent): the flows of the product nodes' children are written to node_flows
//   kind 3  modify pass of a sum layer partition: node_flows = log(flow) - mar
//
// One workgroup streams every parameter of the run, so its time is set by how many bytes it
// keeps in flight.  Eight wavefronts; an operation's index rows are staged in LDS once; every
// phase issues all its loads before consuming any; a sum operation first stages the MFMA B
// operand of all its blocks in LDS (exp(child - column max), or exp(log(flow / mar) - column
// max) for the element flows), then each wavefront runs one pipeline over all its 32-row tiles of
// all blocks, with the parameters (A operand, straight from global memory in the matrix-core
// lane layout) of the next two 64-wide reduction chunks in flight.
#include "pjamd_mfma.h"

using namespace pjamd;

namespace {

constexpr int kThreads = 512;               // eight wavefronts
constexpr int kWaves = kThreads / 64;
constexpr int kCols = 32;                   // batch columns per workgroup
constexpr int kRowGroups = kThreads / kCols;
constexpr int kPad = 64;                    // block reductions are staged in multiples of this
constexpr int kChunk = 32;                  // reduction values per pipelined chunk
constexpr int kPerLane = 16;                // A values per lane per chunk (both operand types)
constexpr int kMaxBlocks = 64;              // node / child blocks per sum operation
constexpr int kQ = 32;                      // B-operand values per thread per pass (16 row groups x 32 = 512)

template <typename T>
struct ChainCap {
    static constexpr int K = PJAMD_CHAIN_MAX_K;   // staged reduction entries of one operation
};
template <>
struct ChainCap<float> {
    static constexpr int K = PJAMD_CHAIN_MAX_K / 2;
};

struct Op {
    int kind, nb, K, bs, cbs, Kc, flags;
    int64_t eoff, a, b, c, d, e;
};

__device__ __forceinline__ Op load_op(const int64_t* __restrict__ o) {
    Op op;
    op.kind = (int)o[0];
    op.nb = (int)o[1];
    op.K = (int)o[2];
    op.bs = (int)o[3];
    op.cbs = (int)o[4];
    op.eoff = o[5];
    op.a = o[6];
    op.b = o[7];
    op.c = o[8];
    op.d = o[9];
    op.e = o[10];
    op.Kc = (int)o[11];
    op.flags = (int)o[12];
    return op;
}

template <typename T>
struct Smem {
    static constexpr int CAP = ChainCap<T>::K;
    static constexpr int LDB = CAP + MF<T>::PAD;
    T Bs[kCols][LDB];                  // MFMA B operand of all blocks: [column][block offset + k]
    int s0[CAP];                       // staged index rows (cids / parids / nids)
    int s1[CAP];                       // staged index rows (pids / parpids / cids)
    int cnt[kMaxBlocks];               // reduction length of each block (children / parent rows)
    int boff[kMaxBlocks + 1];          // block offsets into the B operand (multiples of kPad)
    int first[kMaxBlocks];             // kind 2: first non-zero child code of each block (0: none)
    int nid[kMaxBlocks];               // first node (kind 1) / element (kind 2) row of each block
    float cmax[kMaxBlocks][kCols];
    float red[kRowGroups][kCols];
};

// Column max over the workgroup's row groups into cmax[j].
template <typename T>
__device__ __forceinline__ float column_max(Smem<T>& sm, int j, float m, int rg, int col) {
    sm.red[rg][col] = m;
    __syncthreads();
    if (threadIdx.x < kCols) {
        float v = sm.red[0][threadIdx.x];
#pragma unroll
        for (int g = 1; g < kRowGroups; ++g) v = fmaxf(v, sm.red[g][threadIdx.x]);
        sm.cmax[j][threadIdx.x] = v;
    }
    __syncthreads();
    return sm.cmax[j][col];
}

// Stage n entries of an index table into LDS (callers add the barrier).
__device__ __forceinline__ void stage(int* dst, const int* __restrict__ src, int n) {
    for (int u = threadIdx.x; u < n; u += kThreads) dst[u] = src[u];
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Every load below is unconditional, from an address clamped into the buffer, with the value
// masked afterwards: a load under a per-lane branch makes the compiler wait for every load in
// flight before the next use of any of them.

// kind 0: product partition (gather-add).  A thread owns rows rg + 16 u of its column; U rows x 4
// children of loads are in flight at a time.
template <typename T>
__device__ void chain_prod(const Op& op, const float* __restrict__ nm, float* __restrict__ em,
                           const int* __restrict__ idx, int64_t B, int64_t b0, Smem<T>& sm) {
    const int tid = threadIdx.x, col = tid % kCols, rg = tid / kCols;
    const int64_t b = b0 + col;
    const bool ok = b < B;
    const int64_t bc = ok ? b : B - 1;
    const int lbs = __builtin_ctz((unsigned)op.bs);
    const int rows = op.nb << lbs, K = op.K;
    stage(sm.s0, idx + op.a, op.nb);
    stage(sm.s1, idx + op.b, op.nb * K);
    __syncthreads();
    constexpr int U = 8, KU = 4;
    for (int r0 = rg; r0 < rows; r0 += kRowGroups * U) {
        float acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) acc[u] = 0.0f;
        for (int k0 = 0; k0 < K; k0 += KU) {
            float v[U][KU];
            bool use[U][KU];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int r = r0 + u * kRowGroups;
                const int rr = r < rows ? r : rows - 1;
#pragma unroll
                for (int q = 0; q < KU; ++q) {
                    const int k = k0 + q < K ? k0 + q : K - 1;
                    const int ch = sm.s1[(rr >> lbs) * K + k];
                    v[u][q] = nm[((int64_t)ch + (rr & (op.bs - 1))) * B + bc];
                    use[u][q] = ch != 0 && k0 + q < K;
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int q = 0; q < KU; ++q) acc[u] += use[u][q] ? v[u][q] : 0.0f;
        }
        if (ok) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int r = r0 + u * kRowGroups;
                if (r < rows) em[(op.eoff + sm.s0[r >> lbs] + (r & (op.bs - 1))) * B + b] = acc[u];
            }
        }
    }
}

// kind 3: node_flows[nid + i] = log(flow) - mar for the rows of a sum partition.
template <typename T>
__device__ void chain_modify(const Op& op, float* __restrict__ nf, const float* __restrict__ nm,
                             const int* __restrict__ idx, int64_t B, int64_t b0, Smem<T>& sm) {
    const int tid = threadIdx.x, col = tid % kCols, rg = tid / kCols;
    const int64_t b = b0 + col;
    const int lbs = __builtin_ctz((unsigned)op.bs);
    const int rows = op.nb << lbs;
    stage(sm.s0, idx + op.a, op.nb);
    __syncthreads();
    if (b >= B) return;
    constexpr int U = 8;
    for (int r0 = rg; r0 < rows; r0 += kRowGroups * U) {
        float f[U], m[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int r = r0 + u * kRowGroups;
            const int rr = r < rows ? r : rows - 1;
            const int64_t o = ((int64_t)sm.s0[rr >> lbs] + (rr & (op.bs - 1))) * B + b;
            f[u] = nf[o];
            m[u] = nm[o];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int r = r0 + u * kRowGroups;
            if (r < rows) {
                const int64_t o = ((int64_t)sm.s0[r >> lbs] + (r & (op.bs - 1))) * B + b;
                nf[o] = (m[u] == kNegInf) ? kNegInf : __logf(f[u]) - m[u];
            }
        }
    }
}

// Position of a wavefront's pipeline: tile g (block j = g / tiles per block) and chunk c.
// Wavefront-uniform (scalar registers).
struct Cursor {
    int g, c;
};

// The sum operations (kinds 1 and 2) share the staging and the pipeline; they differ in what the
// B operand holds, where the A operand lives and what the epilogue writes.
//   kind 1: block j = node block; reduction over its n_j children; rows = the block's bs nodes.
//           flags bit 2: every aligned group of 8 children is a run (pids 8-aligned groups step
//           by bs), so a lane's 8 parameters need one index.
//   kind 2: block j = child (product) block; reduction over the np_j * bs parent rows; rows = the
//           block's cbs child rows.  flags bit 2: parameter bases are multiples of 4 (16-byte
//           aligned rows of 8 consecutive parameters).
// Every block gets at least one chunk, so blocks without children / parents produce -inf / 0.
template <typename T, int KIND>
__device__ void chain_sum(const Op& op, float* __restrict__ out, const float* nflows, const float* __restrict__ nm,
                          const float* __restrict__ em, const float* __restrict__ params,
                          const int* __restrict__ idx, int64_t B, int64_t b0, Smem<T>& sm) {
    constexpr int KPL = MF<T>::KPL, KS = 2 * KPL, STEPS = kChunk / KS;
    static_assert(STEPS * KPL == kPerLane, "chunk layout");
    const int tid = threadIdx.x, col = tid % kCols, rg = tid / kCols;
    const int lane = tid & 63, h = lane >> 5, mcol = lane & 31;
    const int w = uni(tid >> 6);
    const int64_t b = b0 + col;
    const bool ok = b < B;
    const int64_t bc = ok ? b : B - 1;
    const int64_t bm = b0 + mcol;
    const int nb = op.nb, K = op.K;
    const int bs = op.bs, lbs = __builtin_ctz((unsigned)bs);
    const bool modified = op.flags & 1, mod_children = op.flags & 2, fast = op.flags & 4;
    const int rows_per_block = KIND == 1 ? bs : op.cbs;

    // ---- stage index rows and block offsets
    stage(sm.s0, idx + op.b, nb * K);
    stage(sm.s1, idx + op.c, nb * K);
    if (tid < nb) {
        sm.cnt[tid] = KIND == 1 ? idx[op.d + tid] : (idx[op.d + tid] << lbs);
        sm.nid[tid] = idx[op.a + tid];
        if (KIND == 2) {
            int f = 0;
            for (int q = op.Kc - 1; q >= 0; --q) {
                const int code = idx[op.e + (int64_t)tid * op.Kc + q];
                if (code != 0) f = code;
            }
            sm.first[tid] = f;
        }
    }
    __syncthreads();
    if (tid == 0) {
        int o = 0;
        for (int j = 0; j < nb; ++j) {
            sm.boff[j] = o;
            const int len = sm.cnt[j];
            o += len > 0 ? (len + kPad - 1) / kPad * kPad : kPad;
        }
        sm.boff[nb] = o;
    }
    __syncthreads();

    // ---- B operand of every block: one pass of loads per block
    // values k = p0 + rg + 16 q of block j: all loads first, then the transform and the mask
    auto bvals = [&](int j, int p0, int len, float* v) {
        float f[kQ], mm[kQ];
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const int k = p0 + rg + kRowGroups * q;
            const int kk = k < len ? k : (len > 0 ? len - 1 : 0);
            if (KIND == 1) {
                f[q] = em[(op.eoff + sm.s0[j * K + kk]) * B + bc];
            } else {
                const int64_t o = ((int64_t)sm.s0[j * K + (kk >> lbs)] + (kk & (bs - 1))) * B + bc;
                f[q] = nflows[o];
                if (!modified) mm[q] = nm[o];
            }
        }
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const int k = p0 + rg + kRowGroups * q;
            float x = f[q];
            if (KIND == 2 && !modified) x = (mm[q] == kNegInf) ? kNegInf : __logf(f[q]) - mm[q];
            v[q] = (k < len && ok) ? x : kNegInf;
        }
    };
    for (int j = 0; j < nb; ++j) {
        const int len = sm.cnt[j];
        const int padded = sm.boff[j + 1] - sm.boff[j];
        float v[kQ];
        float m = kNegInf;
        for (int p0 = 0; p0 < len; p0 += kRowGroups * kQ) {
            bvals(j, p0, len, v);
#pragma unroll
            for (int q = 0; q < kQ; ++q) m = fmaxf(m, v[q]);
        }
        const float mx = column_max(sm, j, m, rg, col);
        const bool one_pass = len <= kRowGroups * kQ;
        for (int p0 = 0; p0 < padded; p0 += kRowGroups * kQ) {
            if (!one_pass) bvals(j, p0, len, v);
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const int k = p0 + rg + kRowGroups * q;
                const float e = (k < len && mx != kNegInf && v[q] != kNegInf) ? __expf(v[q] - mx) : 0.0f;
                if (k < padded) sm.Bs[col][sm.boff[j] + k] = MF<T>::cvt(e);
            }
        }
    }
    __syncthreads();

    // ---- one pipeline per wavefront over (tile, chunk) of all blocks
    const int tpb = rows_per_block >> 5;
    const int ntiles = nb * tpb;
    auto nchunks = [&](int g) {
        const int j = g / tpb;
        return uni((sm.boff[j + 1] - sm.boff[j]) / kChunk);
    };
    auto advance = [&](Cursor u) {
        if (u.g < ntiles && ++u.c >= nchunks(u.g)) {
            u.g += kWaves;
            u.c = 0;
        }
        return u;
    };
    auto load = [&](Cursor u, float* dst) {
        if (u.g >= ntiles) return;
        const int j = u.g / tpb, t = u.g - j * tpb;
        const int row = 32 * t + mcol;
        const int len = uni(sm.cnt[j]);
        const int* s1 = sm.s1 + j * K;
        if (len == 0) {   // a block without children / parents: its one chunk multiplies zeros
#pragma unroll
            for (int i = 0; i < kPerLane; ++i) dst[i] = 0.0f;
            return;
        }
#pragma unroll
        for (int s = 0; s < STEPS; ++s) {
            const int k = u.c * kChunk + s * KS + KPL * h;
            // Reductions beyond the block's length are loaded from clamped (valid, finite)
            // parameters: their B operand is 0, so they add nothing.
            if (KIND == 1) {
                if (KPL > 1 && fast) {
                    // one index per lane: the 8 children are a run (clamped to the last child)
                    const int kc = k < len ? k : ((len - 1) & ~(KPL - 1));
                    const int last = len - 1 - kc;
                    const float* p = params + (int64_t)s1[kc > 0 ? kc : 0] + row;
#pragma unroll
                    for (int q = 0; q < KPL; ++q) dst[s * KPL + q] = p[(int64_t)(q < last ? q : (last > 0 ? last : 0)) * bs];
                } else {
#pragma unroll
                    for (int q = 0; q < KPL; ++q) {
                        const int kk = k + q < len ? k + q : (len > 0 ? len - 1 : 0);
                        dst[s * KPL + q] = params[(int64_t)s1[kk] + row];
                    }
                }
            } else {
                // KPL consecutive parent rows lie in one parent block (KPL divides bs): contiguous
                const int kc = k < len ? k : (len >= KPL ? len - KPL : 0);
                const float* p = params + (int64_t)s1[kc >> lbs] + (int64_t)row * bs + (kc & (bs - 1));
                if (KPL == 8 && fast) {
                    const float4 x0 = reinterpret_cast<const float4*>(p)[0];
                    const float4 x1 = reinterpret_cast<const float4*>(p)[1];
                    dst[s * KPL + 0] = x0.x;
                    dst[s * KPL + 1] = x0.y;
                    dst[s * KPL + 2] = x0.z;
                    dst[s * KPL + 3] = x0.w;
                    dst[s * KPL + 4] = x1.x;
                    dst[s * KPL + 5] = x1.y;
                    dst[s * KPL + 6] = x1.z;
                    dst[s * KPL + 7] = x1.w;
                } else {
#pragma unroll
                    for (int q = 0; q < KPL; ++q) dst[s * KPL + q] = p[q];
                }
            }
        }
    };

    // One chunk of MFMAs from the A values in `cur`, then, at the end of a tile, its epilogue.
    // Stores only: kind 2 parks the raw tile sums in the rows of the block's first child, and the
    // pass after the pipeline turns them into the children's flows.
    floatx16 acc = {0};
    auto compute = [&](Cursor u, const float* cur) {
        const int j = u.g / tpb, t = u.g - j * tpb;
        const int kb = uni(sm.boff[j]) + u.c * kChunk;
#pragma unroll
        for (int s = 0; s < STEPS; ++s) {
            T a[KPL];
#pragma unroll
            for (int q = 0; q < KPL; ++q) a[q] = MF<T>::cvt(cur[s * KPL + q]);
            acc = MF<T>::mma(a, &sm.Bs[mcol][kb + s * KS + KPL * h], acc);
        }
        if (u.c + 1 == nchunks(u.g)) {
            const int target = KIND == 1 ? uni(sm.nid[j]) : uni(sm.first[j] >> 1);
            const bool write = KIND == 1 || uni(sm.first[j]) != 0;
            const float mxm = sm.cmax[j][mcol];
            if (bm < B && write) {
                const int64_t base = (int64_t)target + 32 * t;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const float v = acc[r];
                    float o = v;
                    if (KIND == 1) o = (v > 0.0f && mxm != kNegInf) ? __logf(v) + mxm : kNegInf;
                    out[(base + acc_row(r, lane)) * B + bm] = o;
                }
            }
            acc = floatx16{0};
        }
    };

    // Three rotating register buffers: a chunk's loads are issued three chunks before its MFMAs
    // (at most 48 loads in flight per lane, within the hardware's count of 63), and no in-flight
    // buffer is ever copied (a copy would wait for its loads).
    Cursor ca = {w, 0};
    Cursor cb = advance(ca);
    Cursor cc = advance(cb);
    float bufa[kPerLane], bufb[kPerLane], bufc[kPerLane];
    load(ca, bufa);
    load(cb, bufb);
    load(cc, bufc);
    while (ca.g < ntiles) {
        compute(ca, bufa);
        const Cursor cd = advance(cc);
        load(cd, bufa);
        if (cb.g >= ntiles) break;
        compute(cb, bufb);
        const Cursor ce = advance(cd);
        load(ce, bufb);
        if (cc.g >= ntiles) break;
        compute(cc, bufc);
        const Cursor cf = advance(ce);
        load(cf, bufc);
        ca = cd;
        cb = ce;
        cc = cf;
    }

    if (KIND == 2) {
        // children's flows from the parked tile sums: v = sum * exp(e + column max), stored as
        // log(v) - mar for the marked children
        __syncthreads();
        if (ok) {
            constexpr int U = 8;
            const int R = rows_per_block;
            for (int j = 0; j < nb; ++j) {
                const int first = uni(sm.first[j]);
                if (first == 0) continue;
                const float mx = sm.cmax[j][col];
                const int64_t erow = op.eoff + uni(sm.nid[j]);
                const int64_t park = first >> 1;
                const int* ch = idx + op.e + (int64_t)j * op.Kc;
                for (int r0 = rg; r0 < R; r0 += kRowGroups * U) {
                    float v[U], e[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int r = r0 + u * kRowGroups;
                        const int rr = r < R ? r : R - 1;
                        e[u] = em[(erow + rr) * B + b];
                        v[u] = out[(park + rr) * B + b];
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        v[u] = (mx == kNegInf || e[u] == kNegInf) ? 0.0f : v[u] * __expf(e[u] + mx);
                    for (int q = 0; q < op.Kc; ++q) {
                        const int code = ch[q];
                        if (code == 0) continue;
                        const int64_t cbase = code >> 1;
                        if ((code & 1) && mod_children) {
                            float mv[U];
#pragma unroll
                            for (int u = 0; u < U; ++u) {
                                const int r = r0 + u * kRowGroups;
                                mv[u] = nm[(cbase + (r < R ? r : R - 1)) * B + b];
                            }
#pragma unroll
                            for (int u = 0; u < U; ++u) {
                                const int r = r0 + u * kRowGroups;
                                if (r < R) out[(cbase + r) * B + b] = (mv[u] == kNegInf) ? kNegInf : __logf(v[u]) - mv[u];
                            }
                        } else {
#pragma unroll
                            for (int u = 0; u < U; ++u) {
                                const int r = r0 + u * kRowGroups;
                                if (r < R) out[(cbase + r) * B + b] = v[u];
                            }
                        }
                    }
                }
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void chain_fwd_kernel(float* __restrict__ nm, float* __restrict__ em,
                                                             const float* __restrict__ params,
                                                             const int64_t* __restrict__ ops, int n_ops,
                                                             const int* __restrict__ idx, int64_t B) {
    __shared__ __attribute__((aligned(16))) Smem<T> sm;
    const int64_t b0 = (int64_t)blockIdx.x * kCols;
    for (int o = 0; o < n_ops; ++o) {
        const Op op = load_op(ops + (int64_t)o * PJAMD_CHAIN_OP_FIELDS);
        if (op.kind == 0) {
            chain_prod<T>(op, nm, em, idx, B, b0, sm);
        } else if (op.kind == 1) {
            chain_sum<T, 1>(op, nm, nullptr, nullptr, em, params, idx, B, b0, sm);
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void chain_bwd_kernel(float* nf, const float* __restrict__ nm,
                                                             const float* __restrict__ em,
                                                             const float* __restrict__ params,
                                                             const int64_t* __restrict__ ops, int n_ops,
                                                             const int* __restrict__ idx, int64_t B) {
    __shared__ __attribute__((aligned(16))) Smem<T> sm;
    const int64_t b0 = (int64_t)blockIdx.x * kCols;
    for (int o = 0; o < n_ops; ++o) {
        const Op op = load_op(ops + (int64_t)o * PJAMD_CHAIN_OP_FIELDS);
        if (op.kind == 2) {
            chain_sum<T, 2>(op, nf, nf, nm, em, params, idx, B, b0, sm);
        } else if (op.kind == 3) {
            chain_modify<T>(op, nf, nm, idx, B, b0, sm);
        }
        __syncthreads();
    }
}

}  // namespace

PJAMD_EXPORT int pjamd_chain_forward(float* node_mars, float* element_mars, const float* params,
                                     const int64_t* ops, int n_ops, const int* idx, int64_t B, int precision,
                                     hipStream_t stream) {
    PJAMD_REQUIRE(n_ops >= 0 && B > 0 && precision >= 0 && precision <= 2);
    if (n_ops == 0) return PJAMD_OK;
    PJAMD_REQUIRE(node_mars && element_mars && params && ops && idx);
    const int64_t g = (B + kCols - 1) / kCols;
    PJAMD_REQUIRE(g < ((int64_t)1 << 31));
    if (precision == 1)
        chain_fwd_kernel<__bf16><<<(unsigned)g, kThreads, 0, stream>>>(node_mars, element_mars, params, ops, n_ops,
                                                                       idx, B);
    else if (precision == 2)
        chain_fwd_kernel<_Float16><<<(unsigned)g, kThreads, 0, stream>>>(node_mars, element_mars, params, ops, n_ops,
                                                                         idx, B);
    else
        chain_fwd_kernel<float><<<(unsigned)g, kThreads, 0, stream>>>(node_mars, element_mars, params, ops, n_ops,
                                                                      idx, B);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_chain_backward(float* node_flows, const float* node_mars, const float* element_mars,
                                      const float* params, const int64_t* ops, int n_ops, const int* idx, int64_t B,
                                      int precision, hipStream_t stream) {
    PJAMD_REQUIRE(n_ops >= 0 && B > 0 && precision >= 0 && precision <= 2);
    if (n_ops == 0) return PJAMD_OK;
    PJAMD_REQUIRE(node_flows && node_mars && element_mars && params && ops && idx);
    const int64_t g = (B + kCols - 1) / kCols;
    PJAMD_REQUIRE(g < ((int64_t)1 << 31));
    if (precision == 1)
        chain_bwd_kernel<__bf16><<<(unsigned)g, kThreads, 0, stream>>>(node_flows, node_mars, element_mars, params,
                                                                       ops, n_ops, idx, B);
    else if (precision == 2)
        chain_bwd_kernel<_Float16><<<(unsigned)g, kThreads, 0, stream>>>(node_flows, node_mars, element_mars, params,
                                                                         ops, n_ops, idx, B);
    else
        chain_bwd_kernel<float><<<(unsigned)g, kThreads, 0, stream>>>(node_flows, node_mars, element_mars, params,
                                                                      ops, n_ops, idx, B);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}
