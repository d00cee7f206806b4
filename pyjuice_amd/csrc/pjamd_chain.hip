// Runs of narrow levels in one launch per pass (planner: pyjuice_amd/model/chain.py).
//
// A run is split into components (subtrees that never exchange data inside the run) and the
// batch into slices of 16 columns; workgroup w owns component w % C and slice w / C and walks
// the component's group operations in order.  No workgroup ever waits for another.
//
// A group operation is a dense block of a sum layer: R <= 256 parent rows (node blocks of bs
// rows) over K <= 256 child element rows (product blocks of cbs >= 32 rows, each the sum of at
// most 4 children's node rows).  The matrix-core A operand (theta forward, theta^T backward)
// comes from a packed BF16 / FP16 copy in the 16x16x32 lane layout (pjamd_chain_pack), one
// 16-byte load per lane per instruction; the B operand (exp of child rows, or of parent
// log(flow) - mar) is staged in LDS.  Eight wavefronts; wavefront w owns the 16-row output
// tiles w and w + 8.  Element-wise work is spread as 2 rows x 4 columns per thread (row group
// g = tid / 4, column quad q = tid % 4; a wavefront holds 32 consecutive rows).
//
// Hand-over between operations: the outputs of the last two operations stay in an LDS ring
// (node rows forward, element flows backward) and the planner points the next operations' reads
// there; every other row an operation reads was written before the launch (the planner keeps
// no run that would read rows of its own component from global memory).  Barriers between
// operations therefore cover LDS only and global stores stay in flight.  Software pipeline: the
// loads of operation o + 1 (packed operands, rows computed before the launch) are issued at the
// start of operation o, into a second register set.
//
// The index tables of a component (PJAMD_CHAIN_TAB_INTS entries at most) and its operation
// headers are staged in LDS once, at kernel start.
#include "pjamd_mfma.h"

#include <type_traits>

using namespace pjamd;

namespace {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
constexpr int kCols = PJAMD_CHAIN_COLS;              // batch columns per workgroup (16)
constexpr int kOpw = PJAMD_CHAIN_OP_INTS;
constexpr int kMaxOps = PJAMD_CHAIN_MAX_OPS;
constexpr int kTab = PJAMD_CHAIN_TAB_INTS;
constexpr int kMaxRows = PJAMD_CHAIN_MAX_ROWS;       // R and K bound
constexpr int kMaxKc = PJAMD_CHAIN_MAX_KC;
constexpr int kSteps = kMaxRows / 32;                // 16x16x32 instructions of one tile
constexpr int kTiles = kMaxRows / 16 / kWaves;       // 16-row tiles per wavefront
constexpr int kLd = kMaxRows + 8;                    // LDS row stride of the B operand (16-bit entries)
constexpr int kRing = kCols + 4;                     // LDS row stride of the ring (floats)

static_assert(kCols == 16 && kThreads / 4 * 2 == kMaxRows, "thread layout: 2 rows x 4 columns per thread");

enum { F_R = 0, F_K, F_LBS, F_LCBS, F_NPB, F_NCB, F_KC, F_NID, F_CB, F_A, F_FLAGS };
enum { PAR_LFM = 1, PAR_WRITE = 2, CH_MOD = 4 };

template <typename T>
struct Smem {
    int ops[kMaxOps * kOpw];
    int tab[kTab];
    __attribute__((aligned(16))) T Bs[kCols][kLd];            // B operand: [column][reduction index]
    __attribute__((aligned(16))) float red[kWaves][kCols];    // per wavefront column maxima
    // outputs of the last two operations (slot = operation index & 1): forward node rows by
    // output position, backward element flows by child row
    __attribute__((aligned(16))) float ring[2][kMaxRows][kRing];
};

// Barrier over the workgroup's LDS traffic only: global stores stay in flight.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Stage the component's operation headers and tables; returns the number of operations.
template <typename T>
__device__ __forceinline__ int stage_component(Smem<T>& sm, const int* __restrict__ comp,
                                               const int* __restrict__ ops, const int* __restrict__ tabs, int c) {
    // the planner keeps a component within the LDS capacity; the clamps only bound the copies
    const int op0 = comp[c * PJAMD_CHAIN_COMP_INTS + 0];
    const int nops = min(comp[c * PJAMD_CHAIN_COMP_INTS + 1], kMaxOps);
    const int tab0 = comp[c * PJAMD_CHAIN_COMP_INTS + 2];
    const int ntab = min(comp[c * PJAMD_CHAIN_COMP_INTS + 3], kTab);
    for (int i = threadIdx.x; i < nops * kOpw; i += kThreads) sm.ops[i] = ops[(int64_t)op0 * kOpw + i];
    for (int i = threadIdx.x; i < ntab; i += kThreads) sm.tab[i] = tabs[tab0 + i];
    __syncthreads();
    return nops;
}

struct Hdr {
    int R, K, lbs, lcbs, npb, ncb, KC, nid_off, cb_off, a0, flags;
};

template <typename T>
__device__ __forceinline__ Hdr header(const Smem<T>& sm, int o) {
    const int* op = sm.ops + o * kOpw;
    Hdr h;
    h.R = uni(op[F_R]);
    h.K = uni(op[F_K]);
    h.lbs = uni(op[F_LBS]);
    h.lcbs = uni(op[F_LCBS]);
    h.npb = uni(op[F_NPB]);
    h.ncb = uni(op[F_NCB]);
    h.KC = uni(op[F_KC]);
    h.nid_off = uni(op[F_NID]);
    h.cb_off = uni(op[F_CB]);
    h.a0 = uni(op[F_A]);
    h.flags = uni(op[F_FLAGS]);
    return h;
}

__device__ __forceinline__ void add4(float4& a, const float4 b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    a.w += b.w;
}

__device__ __forceinline__ float4 ninf4() { return make_float4(kNegInf, kNegInf, kNegInf, kNegInf); }

// log(f) - m, element-wise (-inf where m is -inf).
__device__ __forceinline__ float4 lfdm4(const float4 f, const float4 m) {
    float4 y;
    y.x = m.x == kNegInf ? kNegInf : __logf(f.x) - m.x;
    y.y = m.y == kNegInf ? kNegInf : __logf(f.y) - m.y;
    y.z = m.z == kNegInf ? kNegInf : __logf(f.z) - m.z;
    y.w = m.w == kNegInf ? kNegInf : __logf(f.w) - m.w;
    return y;
}

// Column maxima of a [rows, 16] block held as v[2 rows] per thread: reduced across the
// wavefront's 16 row groups with lane exchanges, then across the nw wavefronts that hold rows
// through LDS (one barrier).  m[4]: this thread's quad's maxima; mc: the maximum of column mcol.
template <typename T>
__device__ __forceinline__ void column_max(Smem<T>& sm, const float4 (&v)[2], int q, int nw, int mcol,
                                           float (&m)[4], float& mc) {
    float a[4] = {fmaxf(v[0].x, v[1].x), fmaxf(v[0].y, v[1].y), fmaxf(v[0].z, v[1].z), fmaxf(v[0].w, v[1].w)};
#pragma unroll
    for (int o = 4; o < 64; o <<= 1) {
#pragma unroll
        for (int c = 0; c < 4; ++c) a[c] = fmaxf(a[c], __shfl_xor(a[c], o, 64));
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane < 4) *reinterpret_cast<float4*>(&sm.red[w][4 * q]) = make_float4(a[0], a[1], a[2], a[3]);
    lds_barrier();
    float4 x = *reinterpret_cast<const float4*>(&sm.red[0][4 * q]);
    float y = sm.red[0][mcol];
    for (int i = 1; i < nw; ++i) {
        const float4 z = *reinterpret_cast<const float4*>(&sm.red[i][4 * q]);
        x.x = fmaxf(x.x, z.x);
        x.y = fmaxf(x.y, z.y);
        x.z = fmaxf(x.z, z.z);
        x.w = fmaxf(x.w, z.w);
        y = fmaxf(y, sm.red[i][mcol]);
    }
    m[0] = x.x;
    m[1] = x.y;
    m[2] = x.z;
    m[3] = x.w;
    mc = y;
}

// Bs[4q + c][2g + r] = exp(v[r].c - m[c]) (0 where m is -inf).
template <typename T>
__device__ __forceinline__ void stage_b(Smem<T>& sm, const float4 (&v)[2], const float (&m)[4], int g, int q) {
    typedef T T2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        T2 x;
        const float ns = neg_shift2(m[c]);
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const float e = c == 0 ? v[r].x : c == 1 ? v[r].y : c == 2 ? v[r].z : v[r].w;
            x[r] = (T)op_exp2(fmaf(e, kLog2e, ns));
        }
        *reinterpret_cast<T2*>(&sm.Bs[4 * q + c][2 * g]) = x;
    }
}

// This wavefront's A operands: tiles w and w + 8, n32 instructions each.  Every load is issued
// (branch-free: a tile at or beyond ntile, or a step at or beyond n32, reloads the operation's
// first unit and is never used), so that the waits for them can be counted exactly.
template <typename T8>
__device__ __forceinline__ void load_a(T8 (&a)[kTiles][kSteps], const T8* __restrict__ pa, int a0, int w, int ntile,
                                       int n32, int lane) {
#pragma unroll
    for (int j = 0; j < kTiles; ++j) {
        const int t = w + kWaves * j;
#pragma unroll
        for (int s = 0; s < kSteps; ++s) {
            const int unit = (t < ntile && s < n32) ? (t * n32 + s) * 64 : 0;
            a[j][s] = pa[(int64_t)a0 + unit + lane];
        }
    }
}

// One 16 x 16 output tile: sum over n32 instructions of A x Bs, as two independent accumulation
// chains so that consecutive matrix-core instructions do not wait for each other.
template <typename T>
__device__ __forceinline__ floatx4 tile_mma(const Smem<T>& sm, const typename MF16<T>::vec (&a)[kSteps], int n32,
                                            int lane) {
    typedef typename MF16<T>::vec T8;
    floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    const int col = lane & 15, kq = 8 * (lane >> 4);
#pragma unroll
    for (int s = 0; s < kSteps; s += 2) {
        if (s < n32) acc0 = MF16<T>::mma(a[s], *reinterpret_cast<const T8*>(&sm.Bs[col][32 * s + kq]), acc0);
        if (s + 1 < n32) acc1 = MF16<T>::mma(a[s + 1], *reinterpret_cast<const T8*>(&sm.Bs[col][32 * s + 32 + kq]), acc1);
    }
    return acc0 + acc1;
}

// The operation loop runs two operations per iteration with two register sets, so that the
// independent loads of operation o + 1 are issued at the start of operation o (a whole
// operation ahead of their use).
template <typename T>
__global__ __launch_bounds__(kThreads) void chain_fwd_kernel(float* nm, float* em, const typename MF16<T>::vec* __restrict__ pa,
                                                             const int* __restrict__ comp,
                                                             const int* __restrict__ ops,
                                                             const int* __restrict__ tabs, int C, int64_t B) {
    typedef typename MF16<T>::vec T8;
    __shared__ Smem<T> sm;
    const int c = blockIdx.x % C;
    const int64_t b0 = (int64_t)(blockIdx.x / C) * kCols;
    const int nops = stage_component(sm, comp, ops, tabs, c);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int g = tid >> 2, q = tid & 3;
    const int64_t cq = b0 + 4 * q;
    const bool okq = cq < B;                       // B % 4 == 0: a quad is all in or all out
    const int64_t cqc = okq ? cq : B - 4;
    const int mcol = lane & 15;

    // Product rows 2g, 2g + 1 of an operation: the rows of their first two children (the planner
    // lists children in global memory first), loaded unconditionally and kept raw until use so
    // that nothing waits for them at issue; a ring child or no child reads the zero dummy rows
    // (rows 0 .. cbs - 1 of node_mars).
    auto global_children = [&](const Hdr& d, float4 (&gs)[2][2]) {
        const int k = 2 * g < d.K ? 2 * g : 0;
        const int ent = d.cb_off + (k >> d.lcbs) * (1 + d.KC);
        const int off = k & ((1 << d.lcbs) - 1);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int ref = u < d.KC ? max(uni(sm.tab[ent + 1 + u]), 0) : 0;
#pragma unroll
            for (int r = 0; r < 2; ++r) gs[u][r] = *reinterpret_cast<const float4*>(nm + ((int64_t)ref + off + r) * B + cqc);
        }
    };
    auto prefetch = [&](int o, T8 (&a)[kTiles][kSteps], float4 (&gs)[2][2]) {
        const Hdr d = header(sm, o);
        load_a(a, pa, d.a0, w, d.R >> 4, d.K >> 5, lane);
        global_children(d, gs);
    };
    // FULL: a 256 x 256 group (every bound at its maximum: no guards, constant trip counts)
    auto body = [&](auto full, int o, T8 (&a)[kTiles][kSteps], float4 (&gs)[2][2]) {
        constexpr bool FULL = decltype(full)::value;
        const Hdr d = header(sm, o);
        const int n32 = FULL ? kSteps : d.K >> 5, ntile = FULL ? kTiles * kWaves : d.R >> 4;
        const bool rok = FULL || 2 * g < d.K;
        const int k = rok ? 2 * g : 0;
        const int ent = d.cb_off + (k >> d.lcbs) * (1 + d.KC);
        const int off = k & ((1 << d.lcbs) - 1);
        float4 v[2] = {gs[0][0], gs[0][1]};
        add4(v[0], gs[1][0]);
        add4(v[1], gs[1][1]);
#pragma unroll
        for (int u = 0; u < kMaxKc; ++u) {
            if (u < d.KC) {
                const int ref = uni(sm.tab[ent + 1 + u]);
                if (ref < 0) {
                    // rows of an earlier operation's output: ring slot, first position
                    const int slot = (-ref - 1) >> 8, p0 = ((-ref - 1) & 255) + off;
#pragma unroll
                    for (int r = 0; r < 2; ++r) add4(v[r], *reinterpret_cast<const float4*>(&sm.ring[slot][p0 + r][4 * q]));
                } else if (u >= 2 && ref > 0) {
                    // a third or fourth child in global memory (not prefetched)
#pragma unroll
                    for (int r = 0; r < 2; ++r)
                        add4(v[r], *reinterpret_cast<const float4*>(nm + ((int64_t)ref + off + r) * B + cqc));
                }
            }
        }
        float4 x[2];
#pragma unroll
        for (int r = 0; r < 2; ++r) x[r] = (rok && okq) ? v[r] : ninf4();
        float m[4], mm;
        column_max(sm, x, q, FULL ? kWaves : d.K >> 5, mcol, m, mm);
        if (rok) stage_b(sm, x, m, g, q);
        lds_barrier();
        float (*slot)[kRing] = sm.ring[o & 1];
#pragma unroll
        for (int j = 0; j < kTiles; ++j) {
            const int t = w + kWaves * j;
            if (t < ntile) {
                const floatx4 acc = tile_mma(sm, a[j], n32, lane);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float y = acc[i];
                    slot[16 * t + 4 * (lane >> 4) + i][mcol] = (y > 0.0f && mm != kNegInf) ? __logf(y) + mm : kNegInf;
                }
            }
        }
        lds_barrier();
        // stores: the element rows, and node rows 2g, 2g + 1 (one parent block, bs >= 8) from the
        // ring, 16 bytes per row
        if (rok && okq) {
            const int64_t erow = (int64_t)sm.tab[ent] + off;
#pragma unroll
            for (int r = 0; r < 2; ++r) *reinterpret_cast<float4*>(em + (erow + r) * B + cq) = v[r];
        }
        if ((FULL || 2 * g < d.R) && okq) {
            const int p = 2 * g;
            const int64_t row = (int64_t)sm.tab[d.nid_off + (p >> d.lbs)] + (p & ((1 << d.lbs) - 1));
#pragma unroll
            for (int r = 0; r < 2; ++r)
                *reinterpret_cast<float4*>(nm + (row + r) * B + cq) = *reinterpret_cast<const float4*>(&slot[p + r][4 * q]);
        }
    };

    auto run_op = [&](int o, T8 (&a)[kTiles][kSteps], float4 (&gs)[2][2]) {
        const Hdr d = header(sm, o);
        if (d.K == kMaxRows && d.R == kMaxRows) body(std::true_type{}, o, a, gs);
        else body(std::false_type{}, o, a, gs);
    };

    T8 a0[kTiles][kSteps], a1[kTiles][kSteps];
    float4 gs0[2][2], gs1[2][2];
    if (nops > 0) prefetch(0, a0, gs0);
    for (int o = 0; o < nops; o += 2) {
        if (o + 1 < nops) prefetch(o + 1, a1, gs1);
        run_op(o, a0, gs0);
        if (o + 1 >= nops) break;
        if (o + 2 < nops) prefetch(o + 2, a0, gs0);
        run_op(o + 1, a1, gs1);
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void chain_bwd_kernel(float* nf, const float* __restrict__ nm,
                                                             const float* __restrict__ em,
                                                             const typename MF16<T>::vec* __restrict__ pa,
                                                             const int* __restrict__ comp,
                                                             const int* __restrict__ ops,
                                                             const int* __restrict__ tabs, int C, int64_t B) {
    typedef typename MF16<T>::vec T8;
    __shared__ Smem<T> sm;
    const int c = blockIdx.x % C;
    const int64_t b0 = (int64_t)(blockIdx.x / C) * kCols;
    const int nops = stage_component(sm, comp, ops, tabs, c);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int g = tid >> 2, q = tid & 3;
    const int64_t cq = b0 + 4 * q;
    const bool okq = cq < B;
    const int64_t cqc = okq ? cq : B - 4;
    const int mcol = lane & 15;

    // Loads of one operation that do not depend on the run: the A operands, the element rows of
    // child rows 2g, 2g + 1, the parents' marginals and (flows from before the launch) flows.
    struct Pre {
        T8 a[kTiles][kSteps];
        float4 e[2], mr[2], gf[2];
    };
    // parent rows 2g, 2g + 1 of an operation (inside one parent block) and their flow source:
    // >= 0 a ring row (256 slot + first row; uniform per wavefront, bs >= 32), -1 node_flows
    auto parent_row = [&](const Hdr& d, int64_t& prow, int& src) {
        const int p = 2 * g < d.R ? 2 * g : 0;
        const int pb = p >> d.lbs;
        prow = (int64_t)sm.tab[d.nid_off + pb] + (p & ((1 << d.lbs) - 1));
        src = sm.tab[d.nid_off + d.npb + pb];
    };
    auto load_flows = [&](int64_t prow, float4 (&gf)[2]) {
#pragma unroll
        for (int r = 0; r < 2; ++r) gf[r] = *reinterpret_cast<const float4*>(nf + (prow + r) * B + cqc);
    };
    auto prefetch = [&](int o, Pre& P) {
        const Hdr d = header(sm, o);
        load_a(P.a, pa, d.a0, w, d.K >> 4, d.R >> 5, lane);
        const int k = 2 * g < d.K ? 2 * g : 0;
        const int64_t erow = (int64_t)sm.tab[d.cb_off + (k >> d.lcbs) * (1 + d.KC)] + (k & ((1 << d.lcbs) - 1));
#pragma unroll
        for (int r = 0; r < 2; ++r) P.e[r] = *reinterpret_cast<const float4*>(em + (erow + r) * B + cqc);
        int64_t prow;
        int src;
        parent_row(d, prow, src);
        // the marginals unconditionally (nearly every operation uses them); the flows only where
        // they come from global memory (src is uniform per wavefront)
#pragma unroll
        for (int r = 0; r < 2; ++r) P.mr[r] = *reinterpret_cast<const float4*>(nm + (prow + r) * B + cqc);
        if (src < 0) load_flows(prow, P.gf);
    };
    // (no 256 x 256 specialization as in the forward: with it the kernel needs more than 256
    // VGPRs and spills)
    auto body = [&](int o, Pre& P) {
        const Hdr d = header(sm, o);
        const int n32 = d.R >> 5, ntile = d.K >> 4;
        const int cbs = 1 << d.lcbs;
        int64_t prow;
        int src;
        parent_row(d, prow, src);
        const bool rok = 2 * g < d.R;
        // log(flow) - mar of parent rows 2g, 2g + 1, the flows from the ring (raw, a child block
        // of an earlier operation) or from node_flows
        float4 v[2];
        if (src >= 0) {
            const int rs = src >> 8, k0 = (src & 255) + ((2 * g) & ((1 << d.lbs) - 1));
#pragma unroll
            for (int r = 0; r < 2; ++r) v[r] = *reinterpret_cast<const float4*>(&sm.ring[rs][k0 + r][4 * q]);
        } else {
            v[0] = P.gf[0];
            v[1] = P.gf[1];
        }
        const bool conv = src >= 0 || !(d.flags & PAR_LFM);
        if (conv) {
#pragma unroll
            for (int r = 0; r < 2; ++r) v[r] = lfdm4(v[r], P.mr[r]);
        }
        // modified flows go back to node_flows: raw flows of a modified layer (PAR_WRITE), and
        // ring-fed parents, whose producer stored them raw, whenever the pass modifies
        const bool write_back = conv && (d.flags & (PAR_WRITE | (src >= 0 ? PAR_LFM : 0)));
        float4 x[2];
#pragma unroll
        for (int r = 0; r < 2; ++r) x[r] = (rok && okq) ? v[r] : ninf4();
        float m[4], mm;
        column_max(sm, x, q, d.R >> 5, mcol, m, mm);
        if (rok) stage_b(sm, x, m, g, q);
        lds_barrier();
        float (*slot)[kRing] = sm.ring[o & 1];
#pragma unroll
        for (int j = 0; j < kTiles; ++j) {
            const int t = w + kWaves * j;
            if (t < ntile) {
                const floatx4 acc = tile_mma(sm, P.a[j], n32, lane);
#pragma unroll
                for (int i = 0; i < 4; ++i) slot[16 * t + 4 * (lane >> 4) + i][mcol] = acc[i];
            }
        }
        lds_barrier();
        // child rows 2g, 2g + 1: flow = sum x exp(element row + column max), kept raw in the ring
        const bool kok = 2 * g < d.K;
        const int k = kok ? 2 * g : 0;
        const int ent = d.cb_off + (k >> d.lcbs) * (1 + d.KC);
        const int koff = k & (cbs - 1);
        float4 f[2];
        if (kok) {
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const float4 s4 = *reinterpret_cast<const float4*>(&slot[k + r][4 * q]);
                const float4 e = P.e[r];
                f[r].x = (m[0] == kNegInf || e.x == kNegInf) ? 0.0f : s4.x * __expf(e.x + m[0]);
                f[r].y = (m[1] == kNegInf || e.y == kNegInf) ? 0.0f : s4.y * __expf(e.y + m[1]);
                f[r].z = (m[2] == kNegInf || e.z == kNegInf) ? 0.0f : s4.z * __expf(e.z + m[2]);
                f[r].w = (m[3] == kNegInf || e.w == kNegInf) ? 0.0f : s4.w * __expf(e.w + m[3]);
                *reinterpret_cast<float4*>(&slot[k + r][4 * q]) = f[r];
            }
        }
        lds_barrier();     // ring slot handed over
        // stores
        if (write_back && rok && okq) {
#pragma unroll
            for (int r = 0; r < 2; ++r) *reinterpret_cast<float4*>(nf + (prow + r) * B + cq) = v[r];
        }
        if (kok && okq) {
            // product children (one parent each): code = node row << 2 | marked | ring-fed; a
            // marked child read by no later operation of the run gets log(flow) - mar
#pragma unroll
            for (int u = 0; u < kMaxKc; ++u) {
                const int code = u < d.KC ? uni(sm.tab[ent + 1 + u]) : 0;
                if (code == 0) continue;
                const int64_t cr = (int64_t)(code >> 2) + koff;
                if ((code & 1) && !(code & 2) && (d.flags & CH_MOD)) {
                    float4 mc[2];
#pragma unroll
                    for (int r = 0; r < 2; ++r) mc[r] = *reinterpret_cast<const float4*>(nm + (cr + r) * B + cqc);
#pragma unroll
                    for (int r = 0; r < 2; ++r) *reinterpret_cast<float4*>(nf + (cr + r) * B + cq) = lfdm4(f[r], mc[r]);
                } else {
#pragma unroll
                    for (int r = 0; r < 2; ++r) *reinterpret_cast<float4*>(nf + (cr + r) * B + cq) = f[r];
                }
            }
        }
    };

    Pre P0, P1;
    if (nops > 0) prefetch(0, P0);
    for (int o = 0; o < nops; o += 2) {
        if (o + 1 < nops) prefetch(o + 1, P1);
        body(o, P0);
        if (o + 1 >= nops) break;
        if (o + 2 < nops) prefetch(o + 2, P0);
        body(o + 1, P1);
    }
}

// Packed A operands: workgroup i packs 128 units (16 bytes each) of group wg[3 i], direction
// wg[3 i + 1] (0 forward, 1 backward), from unit wg[3 i + 2].
//   forward  unit (t, s, lane) = theta[16 t + lane % 16][32 s + 8 (lane / 16) + q], q < 8
//   backward unit (t, s, lane) = theta[32 s + 8 (lane / 16) + q][16 t + lane % 16]
// with theta[p][c] = params[pbase[p / bs][c / cbs] + (c % cbs) bs + p % bs].
template <typename T>
__global__ __launch_bounds__(128) void chain_pack_kernel(typename MF16<T>::vec* __restrict__ out,
                                                         const float* __restrict__ params,
                                                         const int64_t* __restrict__ desc,
                                                         const int64_t* __restrict__ pbase,
                                                         const int* __restrict__ wg) {
    typedef typename MF16<T>::vec T8;
    const int gi = wg[3 * blockIdx.x], dir = wg[3 * blockIdx.x + 1];
    const int64_t u = (int64_t)wg[3 * blockIdx.x + 2] + threadIdx.x;
    const int64_t* d = desc + (int64_t)gi * PJAMD_CHAIN_DESC_INTS;
    const int R = (int)d[0], K = (int)d[1], lbs = (int)d[2], lcbs = (int)d[3], ncb = (int)d[5];
    const int bs = 1 << lbs, cbs = 1 << lcbs;
    const int64_t* pb = pbase + d[8];
    const int lane = (int)(u & 63);
    const int64_t m = u >> 6;
    T8 x;
    if (dir == 0) {
        const int n32 = K >> 5;
        const int t = (int)(m / n32), s = (int)(m % n32);
        const int pr = 16 * t + (lane & 15), c0 = 32 * s + 8 * (lane >> 4);
        const int64_t src = pb[(int64_t)(pr >> lbs) * ncb + (c0 >> lcbs)] + (int64_t)(c0 & (cbs - 1)) * bs + (pr & (bs - 1));
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = (T)params[src + (int64_t)i * bs];
        out[d[6] + u] = x;
    } else {
        const int n32 = R >> 5;
        const int t = (int)(m / n32), s = (int)(m % n32);
        const int ch = 16 * t + (lane & 15), p0 = 32 * s + 8 * (lane >> 4);
        const int64_t src = pb[(int64_t)(p0 >> lbs) * ncb + (ch >> lcbs)] + (int64_t)(ch & (cbs - 1)) * bs + (p0 & (bs - 1));
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = (T)params[src + i];
        out[d[7] + u] = x;
    }
}

// Lane-layout check of the 16x16x32 instructions: C[16][16] = A[16][32] B[32][16].
template <typename T>
__global__ void mfma16_layout_probe(const float* A, const float* B, float* C) {
    typedef typename MF16<T>::vec T8;
    __shared__ __attribute__((aligned(16))) T As[16][32];
    __shared__ __attribute__((aligned(16))) T Bt[16][32];
    const int lane = threadIdx.x;
    for (int i = lane; i < 16 * 32; i += 64) {
        const int r = i / 32, k = i % 32;
        As[r][k] = (T)A[r * 32 + k];
        Bt[r][k] = (T)B[k * 16 + r];
    }
    __syncthreads();
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = MF16<T>::mma(*reinterpret_cast<const T8*>(&As[lane & 15][8 * (lane >> 4)]),
                       *reinterpret_cast<const T8*>(&Bt[lane & 15][8 * (lane >> 4)]), acc);
    for (int i = 0; i < 4; ++i) C[(4 * (lane >> 4) + i) * 16 + (lane & 15)] = acc[i];
}

}  // namespace

PJAMD_EXPORT int pjamd_chain_pack(void* packed, const float* params, const int64_t* desc, const int64_t* pbase,
                                  const int* wg, int n_wg, int precision, pjamd_stream_t stream) {
    PJAMD_REQUIRE(n_wg >= 0 && (precision == 1 || precision == 2));
    if (n_wg == 0) return PJAMD_OK;
    PJAMD_REQUIRE(packed && params && desc && pbase && wg);
    if (precision == 1)
        chain_pack_kernel<__bf16><<<n_wg, 128, 0, stream>>>(reinterpret_cast<bf16x8*>(packed), params, desc, pbase, wg);
    else
        chain_pack_kernel<_Float16><<<n_wg, 128, 0, stream>>>(reinterpret_cast<halfx8*>(packed), params, desc, pbase,
                                                              wg);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_chain_forward(float* node_mars, float* element_mars, const void* packed, const int* comp,
                                     const int* ops, const int* tabs, int n_comp, int64_t B, int precision,
                                     pjamd_stream_t stream) {
    PJAMD_REQUIRE(n_comp >= 0 && B > 0 && B % 4 == 0 && (precision == 1 || precision == 2));
    if (n_comp == 0) return PJAMD_OK;
    PJAMD_REQUIRE(node_mars && element_mars && packed && comp && ops && tabs);
    const int64_t g = (B + kCols - 1) / kCols * n_comp;
    PJAMD_REQUIRE(g < ((int64_t)1 << 31));
    if (precision == 1)
        chain_fwd_kernel<__bf16><<<(unsigned)g, kThreads, 0, stream>>>(
            node_mars, element_mars, reinterpret_cast<const bf16x8*>(packed), comp, ops, tabs, n_comp, B);
    else
        chain_fwd_kernel<_Float16><<<(unsigned)g, kThreads, 0, stream>>>(
            node_mars, element_mars, reinterpret_cast<const halfx8*>(packed), comp, ops, tabs, n_comp, B);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_chain_backward(float* node_flows, const float* node_mars, const float* element_mars,
                                      const void* packed, const int* comp, const int* ops, const int* tabs, int n_comp,
                                      int64_t B, int precision, pjamd_stream_t stream) {
    PJAMD_REQUIRE(n_comp >= 0 && B > 0 && B % 4 == 0 && (precision == 1 || precision == 2));
    if (n_comp == 0) return PJAMD_OK;
    PJAMD_REQUIRE(node_flows && node_mars && element_mars && packed && comp && ops && tabs);
    const int64_t g = (B + kCols - 1) / kCols * n_comp;
    PJAMD_REQUIRE(g < ((int64_t)1 << 31));
    if (precision == 1)
        chain_bwd_kernel<__bf16><<<(unsigned)g, kThreads, 0, stream>>>(
            node_flows, node_mars, element_mars, reinterpret_cast<const bf16x8*>(packed), comp, ops, tabs, n_comp, B);
    else
        chain_bwd_kernel<_Float16><<<(unsigned)g, kThreads, 0, stream>>>(
            node_flows, node_mars, element_mars, reinterpret_cast<const halfx8*>(packed), comp, ops, tabs, n_comp, B);
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}

PJAMD_EXPORT int pjamd_debug_mfma16_probe(int precision, const float* A, const float* B, float* C,
                                          pjamd_stream_t stream) {
    PJAMD_REQUIRE(A && B && C);
    if (precision == 1) mfma16_layout_probe<__bf16><<<1, 64, 0, stream>>>(A, B, C);
    else if (precision == 2) mfma16_layout_probe<_Float16><<<1, 64, 0, stream>>>(A, B, C);
    else return PJAMD_EINVAL;
    PJAMD_LAUNCH_CHECK();
    return PJAMD_OK;
}
