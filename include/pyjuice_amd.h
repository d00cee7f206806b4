/*
 * pyjuice_amd — C ABI of the MI355X (gfx950) probabilistic-circuit kernels.
 *
 * One shared library, libpyjuice_amd.so, built from the HIP sources in pyjuice_amd/csrc.  Every entry point
 * is a plain C function over device pointers and sizes; it enqueues its kernels on `stream`
 * and returns without synchronising.  Return value: 0 on success, a positive hipError_t when
 * a launch failed, PJAMD_EINVAL (-1) for invalid sizes / argument combinations and
 * PJAMD_EUNSUPPORTED (-2) for a configuration the library does not implement and
 * PJAMD_ETOOLARGE (-3) for a request beyond an allocation-free entry point's fixed workspace
 * (pjamd_error_string() names each code).
 *
 * Memory layout shared by all entry points (device memory, row-major):
 *   node_mars / node_flows        float [num_nodes][B]      one row per circuit node
 *   element_mars / element_flows  float [num_elements][B]   one row per product node
 *   params                        float [num_params]        sum-edge parameters; the bs
 *                                  parameters of a (node block, child) edge are contiguous
 *   param_flows                   float or double [num_param_flows]
 *   data                          int32 [num_vars][B]       (see pjamd_data_to_vb)
 * Node / element row 0 .. (max block size - 1) are the dummy rows the layer compiler reserves:
 * index 0 in any child / parent index table means "padding" and is never written.
 *
 * The index tables (nids, cids, pids, pfids, chids, parids, parpids, n_chs / n_pars) are the
 * partition tensors produced by the layer compiler (pyjuice_amd/layer/compilation.py, which
 * reproduces src/pyjuice/layer/compilation.py of PyJuice bit for bit), as int32.
 *
 * Each declaration cites the PyJuice function whose behaviour it replaces
 * (paths relative to the PyJuice source tree).
 */
#ifndef PYJUICE_AMD_H
#define PYJUICE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same type as hipStream_t; a null stream means the default stream. */
typedef struct ihipStream_t* pjamd_stream_t;

enum {
    PJAMD_STATUS_OK = 0,
    PJAMD_STATUS_EINVAL = -1,
    PJAMD_STATUS_EUNSUPPORTED = -2,
    PJAMD_STATUS_ETOOLARGE = -3      /* beyond the fixed workspace of an allocation-free entry point */
};

/* Sum-layer arithmetic (`precision` arguments). */
enum {
    PJAMD_PREC_FP32_MFMA = 0, /* FP32 operands on v_mfma_f32_32x32x2_f32 (block size >= 32) */
    PJAMD_PREC_BF16_MFMA = 1, /* BF16 operands, FP32 accumulate, per-tile max subtraction      */
    PJAMD_PREC_FP16_MFMA = 2, /* FP16 operands (parameter flows run in BF16), FP32 accumulate  */
    PJAMD_PREC_FP32_VALU = 3  /* FP32 on the vector ALUs, any block size                      */
};
/* Block sizes below 32 always run on the FP32 vector path, whatever `precision` says. */

/* Input distributions (`kind` arguments). */
enum {
    PJAMD_DIST_CATEGORICAL = 0, /* src/pyjuice/nodes/distributions/categorical.py */
    PJAMD_DIST_BERNOULLI = 1    /* src/pyjuice/nodes/distributions/bernoulli.py   */
};

/* ----------------------------------------------------------------------------- library */

/* ABI revision of this header (no PyJuice counterpart). */
int pjamd_abi_version(void);

/* Static text for a status code (no PyJuice counterpart). */
const char* pjamd_error_string(int code);

/* p[i] = log(p[i]) for i < n (-inf where p[i] == 0): hands the node-flow buffer back in log
 * space after a linear-space flow backward.  Replaces the log-space storage of node flows under
 * TensorCircuit.backward(logspace_flows=True) (src/pyjuice/model/tensorcircuit.py:393-417,
 * layer/sum_layer.py:1376-1392). */
int pjamd_flows_to_log(float* p, int64_t n, pjamd_stream_t stream);

/* p[0 .. n) = v.  Replaces the buffer initialisation of TensorCircuit._init_buffer
 * (src/pyjuice/model/tensorcircuit.py:858-878) and the flow / dummy-row fills of
 * TensorCircuit.forward / backward (tensorcircuit.py:207-342, 344-538).  A kernel, not
 * hipMemsetAsync, so it can be recorded into HIP graphs. */
int pjamd_fill_f32(float* p, float v, int64_t n, pjamd_stream_t stream);

/* out[v][b] = (int32) src[b][v] for a [B][V] batch of one-byte (uint8 / bool, src_code 0; int8,
 * src_code 3), int32 (src_code 1) or int64 (src_code 2) values.  Replaces `inputs.permute(1, 0)` of TensorCircuit.forward
 * (tensorcircuit.py:226-230). */
int pjamd_data_to_vb(int* out, int src_code, const void* src, int64_t B, int64_t V, pjamd_stream_t stream);

/* data[v][b] = -1 wherever the missing mask is set: a marginalized entry, which
 * pjamd_input_forward gives log-probability 0 and pjamd_input_missing_flows takes the flows of.
 * data is int32 [V][B] as pjamd_data_to_vb writes it; mask is uint8, [V] (mode 0), [B][V]
 * (mode 1) or [V][B] (mode 2).  Replaces the missing_mask pass of InputLayer.forward
 * (src/pyjuice/layer/input_layer.py:374-401, _fw_missing_mask_kernel 1058-1095) and the
 * mask loads of the flows kernel (input_layer.py:1120-1133). */
int pjamd_data_mask(int* data, const uint8_t* mask, int mode, int64_t B, int64_t V, pjamd_stream_t stream);

/* ----------------------------------------------------------------------------- input layers */

/* node_mars[node_offset + n][b] = log p_n(data[vids[n]][b]) for the num_nodes nodes of one
 * input layer; s_pids[n] is node n's first parameter, ncats[n] its category count and kmax the
 * largest category count of the layer (Categorical, kmax >= every ncats[n]: up to 256 the
 * parameter rows are staged in LDS; ignored for Bernoulli).  A negative data value is a
 * marginalized entry (log p = 0); a category >= ncats[n] has p = 0.
 * Replaces InputLayer.forward (src/pyjuice/layer/input_layer.py:283-404) with
 * Categorical.fw_mar_fn (categorical.py:83-92) / Bernoulli.fw_mar_fn (bernoulli.py:65-71). */
int pjamd_input_forward(int kind, const float* params, float* node_mars, const int* data, const int* vids,
                        const int64_t* s_pids, const int* ncats, int64_t num_nodes, int64_t node_offset, int64_t B,
                        int kmax, pjamd_stream_t stream);

/* pflows[s_pfids[n] + data[vids[n]][b]] += node_flows[node_offset + n][b] summed over b
 * (Categorical; Bernoulli adds the flows of samples with x = 1 to slot 0 and of x = 0 to
 * slot 1).  pf64 selects a double accumulator.  layout = (largest category count of the layer
 * << 3) | 4 if the accumulator is zero (slots are stored, not read; not with 1) | 2 if every
 * tile of 8 consecutive nodes (from node 0) shares one variable and one category count (lets
 * the kernel sort the batch once per tile instead of adding per sample) | 1 if several nodes
 * share parameter-flow slots (then the adds are atomic).
 * Replaces InputLayer.backward
 * (input_layer.py:406-571) with Categorical.bk_flow_fn (categorical.py:94-101) /
 * Bernoulli.bk_flow_fn (bernoulli.py:73-77).  Negative data values (marginalized entries) are
 * skipped; see pjamd_input_missing_flows. */
int pjamd_input_backward(int kind, void* pflows, const float* node_flows, int pf64, const int* data, const int* vids,
                         const int64_t* s_pfids, const int* ncats, int64_t num_nodes, int64_t node_offset, int64_t B,
                         int layout, pjamd_stream_t stream);

/* pjamd_input_backward for a Categorical layer whose parameter-flow ranges are shared by tied
 * vectors: nodes group_nodes[group_offsets[g] .. group_offsets[g + 1]) (indices into the layer,
 * one category count per group, kmax <= 16384) are all the nodes with flow range
 * s_pfids[group_nodes[group_offsets[g]]]; each group's flows are summed over its members and
 * batch, then its slots are written once, without atomics (stored when overwrite != 0, which
 * needs the groups to cover every slot of the accumulator, else added).  Returns
 * PJAMD_EUNSUPPORTED for Bernoulli layers or larger category counts.  Replaces the tied-node
 * part of InputLayer.backward (input_layer.py:406-571): the reference gives up to
 * max_tied_ns_per_parflow_block tied vectors one flow range (input_layer.py:83-110) and adds
 * their flows atomically in Categorical.bk_flow_fn (categorical.py:94-101). */
int pjamd_input_backward_tied(int kind, void* pflows, const float* node_flows, int pf64, const int* data,
                              const int* vids, const int64_t* s_pfids, const int* ncats, const int* group_offsets,
                              const int* group_nodes, int64_t n_groups, int kmax, int64_t node_offset, int64_t B,
                              int overwrite, pjamd_stream_t stream);

/* Flows of the marginalized entries (negative data values) of one input layer, after
 * pjamd_input_backward: Categorical node n adds params[s_pids[n] + k] * (its flows summed over
 * the marginalized batch columns) to pflows[s_pfids[n] + k] for every category k (atomically);
 * Bernoulli nodes take none, as in the reference.  Replaces the masked pass of
 * InputLayer.backward (input_layer.py:529-566) with Categorical.bk_flow_mask_fn
 * (categorical.py:103-134). */
int pjamd_input_missing_flows(int kind, void* pflows, int pf64, const float* node_flows, const float* params,
                              const int* data, const int* vids, const int64_t* s_pids, const int64_t* s_pfids,
                              const int* ncats, int64_t num_nodes, int64_t node_offset, int64_t B,
                              pjamd_stream_t stream);

/* Mini-batch EM of the num_src source nodes src[] of an input layer:
 * theta <- (1 - step) theta + step * (flows + pseudocount / K) / (sum flows + pseudocount),
 * accumulated in double.  Replaces InputLayer.mini_batch_em (input_layer.py:663-700) with
 * Categorical.small_ncats_em_fn / large_ncats_em_fn (categorical.py:164-245) and
 * Bernoulli.em_fn (bernoulli.py:91-101). */
int pjamd_input_em(int kind, float* params, const void* pflows, int pf64, const int* src, const int64_t* s_pids,
                   const int64_t* s_pfids, const int* ncats, int64_t num_src, float step, float pseudocount,
                   int keep_zero, pjamd_stream_t stream);

/* pjamd_input_backward of a Categorical layer with the slots stored (the accumulator counts as
 * zero: every node owns its slots) followed by pjamd_input_em of the same nodes, in one pass:
 * each workgroup holds a node's whole flow histogram, so it stores the slots and updates the
 * node's parameters in place from them (same formula, double accumulation).  Saves the EM
 * launch's re-read of the slots and parameters.  Every node must be its own EM source with its
 * own parameter and slot ranges, every tile of 8 consecutive nodes (from node 0) must share one
 * variable and one category count, and no entry may be marginalized; B % 4 == 0, B <= 2048,
 * kmax <= 1024 and 16-byte aligned rows, else PJAMD_EUNSUPPORTED (the caller then runs the two
 * passes).  No single counterpart in the reference: replaces InputLayer.backward
 * (input_layer.py:406-571) followed by InputLayer.mini_batch_em (input_layer.py:663-700) with
 * Categorical.bk_flow_fn / small_ncats_em_fn (categorical.py:94-101, 164-205) when a training step
 * runs them back to back. */
int pjamd_input_backward_em(void* pflows, const float* node_flows, int pf64, const int* data, const int* vids,
                            const int64_t* s_pfids, const int* ncats, int64_t num_nodes, int64_t node_offset,
                            int64_t B, int kmax, float* params, const int64_t* s_pids, float step,
                            float pseudocount, int keep_zero, pjamd_stream_t stream);

/* For every (node n, sample s) with node_act[node_offset + n][s] != 0, write a value of
 * variable vids[n] and clear the flag: Categorical takes the first category whose FP32 running
 * CDF (left to right) reaches u, or 0 if none does; Bernoulli takes u < theta.  u is
 * uniforms[node_offset + n][s], or, when uniforms is NULL, the counter-based uniform of
 * (node_offset + n, s) under the key seed[0..1] (Philox4x32-10, see pjamd_sample_sum).  out is
 * int64 (out_i64 = 1) or uint8, [num_vars][S].  cdf (optional, Categorical with K <= 4096): the
 * parameters' running sums from pjamd_sample_prefix, used instead of summing per workgroup.
 * Replaces InputLayer.sample
 * (input_layer.py:573-661) with Categorical.sample_fn (categorical.py:137-162) /
 * Bernoulli.sample_fn (bernoulli.py:79-89), which draw tl.rand(seed, offset). */
int pjamd_input_sample(int kind, void* out, const float* params, const float* cdf, uint8_t* node_act,
                       const float* uniforms,
                       const uint32_t* seed, const int* vids, const int64_t* s_pids, const int* ncats, int out_i64,
                       int64_t num_nodes, int64_t node_offset, int64_t S, pjamd_stream_t stream);

/* ----------------------------------------------------------------------------- product layers */

/* out[nids[j] + i][b] (+)= sum_k in[cids[j][k] + i][b] for node block j < n_blocks, row i < bs;
 * cids[j][k] == 0 is padding.  accum = 1 adds to out.  Replaces ProdLayer.forward
 * (src/pyjuice/layer/prod_layer.py:160-204, kernels at 265-555). */
int pjamd_prod_forward(float* out, const float* in, const int* nids, const int* cids, int n_blocks, int K, int bs,
                       int64_t B, int accum, pjamd_stream_t stream);

/* node_flows[u_cids[j] + i][b] += sum_k element_flows[parids[j][k] + i][b]: the backward
 * partitions of a product layer.  Replaces ProdLayer.backward (prod_layer.py:206-261). */
int pjamd_prod_backward(float* node_flows, const float* element_flows, const int* u_cids, const int* parids,
                        int n_blocks, int K, int bs, int64_t B, pjamd_stream_t stream);

/* ----------------------------------------------------------------------------- sum layers */

/* Layout promise bit of pjamd_sum_forward's child_run argument. */
#define PJAMD_PARAMS_ALIGNED (1 << 30)

/* node_mars[nids[j] + i][b] = log sum_{k < n_chs[j]} params[pids[j][k] + i] *
 * exp(element_mars[cids[j][k]][b]) for one forward partition (n_blocks node blocks of bs nodes,
 * index rows of width K).  child_run is a layout promise (0 = none): every child_run-aligned
 * group of index entries names consecutive element rows, cids[j][k0 + i] = cids[j][k0] + i, with
 * parameters bs apart, pids[j][k0 + i] = pids[j][k0] + i * bs; the compiler's index rows keep it
 * for child blocks of that size (PyJuice's block-sparse kernels read them the same way through
 * cids_start / cids_increment).  child_run may carry PJAMD_PARAMS_ALIGNED: every pids entry of
 * the partition's real children is a multiple of 4 (16-byte-aligned parameter rows, read with
 * 16-byte loads when B % 4 == 0 too).  Replaces SumLayer.forward
 * (src/pyjuice/layer/sum_layer.py:220-262) with _forward_block_sparse (sum_layer.py:802-966) and
 * _forward_sparse (1106-1214). */
int pjamd_sum_forward(float* node_mars, const float* element_mars, const float* params, const int* nids,
                      const int* cids, const int* pids, const int* n_chs, int n_blocks, int K, int bs, int64_t B,
                      int child_run, int precision, pjamd_stream_t stream);

/* Launch hints (no PyJuice counterpart), or-ed into pjamd_sum_forward's child_run and the flags
 * of pjamd_sum_backward_ele_to_children: force the preloaded form (the first two reduction stages
 * in flight together) or the plain form of the 64 x 64 16-byte-load kernels, or (forward, block
 * sizes that are multiples of 128, child runs >= 128) the 128 x 128 tile form, instead of the
 * default choice (64 x 64 tiles, preloaded below 256 workgroups).  Set per layer by the launch
 * autotuner. */
#define PJAMD_LAUNCH_PRE (1 << 28)
#define PJAMD_LAUNCH_NOPRE (1 << 27)
#define PJAMD_LAUNCH_WIDE (1 << 26)

/* pjamd_sum_forward as one persistent launch: min(tiles, resident workgroups) workgroups take the
 * (node block, row tile, batch tile) items in order from the counter queue[0] until none is left,
 * so node blocks of uneven fan-in (pruned circuits; the merged table lists the heaviest first)
 * need no dispatch of their own; the last workgroup to finish resets queue[0..1] to zero.  queue:
 * two int32 in device memory, zero before the first call; calls sharing a queue must not overlap.
 * BF16 / FP16 operands, aligned parameter rows with runs >= 16, B % 4 == 0, bs 32 or a multiple of
 * 64, else PJAMD_EUNSUPPORTED.  Same results as pjamd_sum_forward.  Replaces SumLayer.forward's
 * per-partition launches (src/pyjuice/layer/sum_layer.py:220-262). */
int pjamd_sum_forward_queue(float* node_mars, const float* element_mars, const float* params, const int* nids,
                            const int* cids, const int* pids, const int* n_chs, int n_blocks, int K, int bs,
                            int64_t B, int child_run, int* queue, int precision, pjamd_stream_t stream);

/* pjamd_sum_forward with each node block's reduction split over n_splits workgroups (BF16 / FP16
 * operands: precision 1 or 2), for layers of few tiles and long child lists: split s takes an
 * equal share of the 128-child stages and stores its partial sum and column max into workspace,
 * then a second kernel combines the partials into node_mars (their log-sum-exp).  workspace:
 * float [n_splits][n_blocks * bs][B] followed by float [n_splits][n_blocks * bs / 64][B]; calls
 * sharing a workspace must not overlap.  n_blocks * bs < 65536.  Requires child_run with
 * PJAMD_PARAMS_ALIGNED and a run of >= 128, bs a multiple of 64 and B % 4 == 0, else returns
 * PJAMD_EUNSUPPORTED.  Same results as pjamd_sum_forward up to the rounding of the 16-bit
 * operands (each split scales by its own column max).  Replaces SumLayer.forward
 * (src/pyjuice/layer/sum_layer.py:220-262), whose block-sparse kernel runs one program per node
 * tile over the whole child list (sum_layer.py:454-568). */
int pjamd_sum_forward_split(float* node_mars, const float* element_mars, const float* params, const int* nids,
                            const int* cids, const int* pids, const int* n_chs, int n_blocks, int K, int bs,
                            int64_t B, int child_run, int n_splits, float* workspace, int precision,
                            pjamd_stream_t stream);

/* Size in floats of the fixed device-resident workspace of pjamd_sum_forward_split_noalloc
 * (32 MiB per device). */
#define PJAMD_SPLIT_SCRATCH_FLOATS (8 << 20)

/* pjamd_sum_forward_split without a workspace argument, allocating nothing: no counterpart in
 * PyJuice.  For callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The partials live in a fixed array of
 * PJAMD_SPLIT_SCRATCH_FLOATS floats that each device holds until the process ends; a call needing
 * n_splits * (n_blocks * bs + ceil(n_blocks * bs / 64)) * B floats or more than that returns
 * PJAMD_ETOOLARGE before any launch.  The array is looked up on the current device (stream must
 * belong to it) on the first call there.  Every call on a device shares that device's array, so
 * calls that overlap on one device (other streams, other threads) race.  Arguments and results
 * otherwise as pjamd_sum_forward_split, with the same kernels. */
int pjamd_sum_forward_split_noalloc(float* node_mars, const float* element_mars, const float* params,
                                    const int* nids, const int* cids, const int* pids, const int* n_chs,
                                    int n_blocks, int K, int bs, int64_t B, int child_run, int n_splits,
                                    int precision, pjamd_stream_t stream);

/* pjamd_sum_forward with the product layer below it fused in (BF16 / FP16 operands: precision 1
 * or 2).  The partition's element rows are not read from element_mars but summed from the node
 * rows of the products' children, and stored into element_mars on the way (the flow backward
 * reads them).  prod_bases is int32 [n_blocks][ceil(K / 128)][prod_k], prod_k in {2, 4, 8}: for
 * node block j and its stage s of 128 children (element rows cids[j][128 s] + 0..127, a run
 * inside one product block), the first node row of each child block of that product block plus
 * the offset of the stage's first row in the product block; unused entries point into the
 * dummy node rows [0, max element block size), which hold 0.  Requires child_run with
 * PJAMD_PARAMS_ALIGNED and a run of >= 128, bs a multiple of 64 and B % 4 == 0, else returns
 * PJAMD_EUNSUPPORTED (then call pjamd_prod_forward and pjamd_sum_forward).  Replaces
 * ProdLayer.forward (src/pyjuice/layer/prod_layer.py:160-204) followed by SumLayer.forward
 * (sum_layer.py:220-262) for two consecutive layer groups of TensorCircuit.forward
 * (tensorcircuit.py:269-283). */
int pjamd_sum_forward_prod(float* node_mars, float* element_mars, const float* params, const int* nids,
                           const int* cids, const int* pids, const int* n_chs, int n_blocks, int K, int bs, int64_t B,
                           int child_run, const int* prod_bases, int prod_k, int precision, pjamd_stream_t stream);

/* node_flows[r][b] <- log(node_flows[r][b]) - node_mars[r][b] for the rows of the partition's
 * node blocks (-inf where the marginal is -inf).  Replaces SumLayer._bk_triton_modify_flow
 * (sum_layer.py:1355-1515). */
int pjamd_sum_modify_flows(float* node_flows, const float* node_mars, const int* nids, int n_blocks, int bs,
                           int64_t B, pjamd_stream_t stream);

/* element_flows[chids[j] + c][b] (+)= exp(element_mars[..][b]) * sum_{k < n_pars[j], i < bs}
 * params[parpids[j][k] + c * bs + i] * node_flows[parids[j][k] + i][b] / exp(node_mars[..][b])
 * for the n_cblocks child blocks (cbs rows) of one backward partition.  flags bit 0: node flows
 * already hold log(flow) - mar; bit 1: add to element_flows; bit 2: every parpids entry of the
 * real parents is a multiple of 4 (16-byte-aligned parameter rows).  Replaces
 * SumLayer._backward_block_sparse_ele_flows (sum_layer.py:1891-2048) and
 * _backward_sparse_ele_flows (2668-2791). */
int pjamd_sum_backward_ele(float* element_flows, const float* node_flows, const float* node_mars,
                           const float* element_mars, const float* params, const int* chids, const int* parids,
                           const int* parpids, const int* n_pars, int n_cblocks, int KP, int cbs, int bs, int64_t B,
                           int flags, int precision, pjamd_stream_t stream);

/* pjamd_sum_backward_ele fused with the product-layer backward below it: instead of element flows,
 * the flow of each product row is written straight into the node flows of the product's
 * children.  Valid when every child node has that product as its only parent (a product passes
 * its flow unchanged to each child, so no accumulation is needed) and the children are blocks of
 * the product block's size (cbs).  ch_nodes is int32 [n_cblocks][Kc]: for the product block
 * chids[j], 2 * (first node of its k-th child block) + 1 if that child's flows are to be stored as
 * log(flow) - mar, 0 = padding.  flags bit 0: the parents' flows hold log(flow) - mar; bit 1:
 * honour the log(flow) - mar marks of ch_nodes; bit 2: as pjamd_sum_backward_ele.  Replaces
 * SumLayer._backward_block_sparse_ele_flows
 * (sum_layer.py:1891-2048) followed by ProdLayer.backward (prod_layer.py:206-261) and, for the
 * children, SumLayer._bk_triton_modify_flow (sum_layer.py:1439-1515). */
int pjamd_sum_backward_ele_to_children(float* node_flows, const float* node_mars, const float* element_mars,
                                       const float* params, const int* chids, const int* parids, const int* parpids,
                                       const int* n_pars, const int* ch_nodes, int Kc, int n_cblocks, int KP, int cbs,
                                       int bs, int64_t B, int flags, int precision, pjamd_stream_t stream);

/* pjamd_sum_backward_ele_to_children with the reduction over parent rows split over n_splits
 * workgroups (BF16 / FP16 operands, parent blocks of >= 128 rows, child blocks a multiple of 64
 * rows, flags bit 2 set, B % 4 == 0; else PJAMD_EUNSUPPORTED), for layers of few tiles and long
 * parent lists: each split stores its partial sums and column maxima into workspace (laid out as
 * for pjamd_sum_forward_split, over the n_cblocks * cbs child rows, < 65536), and a second kernel
 * combines them and writes the children's flows.  Same results as
 * pjamd_sum_backward_ele_to_children up to the rounding of the 16-bit operands.  Replaces
 * SumLayer._backward_block_sparse_ele_flows (sum_layer.py:1891-2048) followed by
 * ProdLayer.backward (prod_layer.py:206-261). */
int pjamd_sum_backward_ele_split(float* node_flows, const float* node_mars, const float* element_mars,
                                 const float* params, const int* chids, const int* parids, const int* parpids,
                                 const int* n_pars, const int* ch_nodes, int Kc, int n_cblocks, int KP, int cbs, int bs,
                                 int64_t B, int flags, int n_splits, float* workspace, int precision,
                                 pjamd_stream_t stream);

/* param_flows[pfids[j][k] + i] += sum_b params[pids[j][k] + i] * node_flows[nids[j] + i][b] *
 * exp(element_mars[cids[j][k]][b] - node_mars[nids[j] + i][b]) for one forward partition.
 * flags bit 0: flows hold log(flow) - mar; bit 1: slots are shared (atomic adds); bit 2: negate;
 * bit 3: param_flows is double; bit 4: the accumulator is zero and every slot has one writer, so
 * slots are stored instead of read and added (not with bit 1).  Replaces SumLayer._backward_block_sparse_par_flows
 * (sum_layer.py:2284-2412) and _backward_sparse_par_flows (2904-3023). */
int pjamd_sum_backward_par(void* param_flows, const float* node_flows, const float* node_mars,
                           const float* element_mars, const float* params, const int* nids, const int* cids,
                           const int* pids, const int* pfids, const int* n_chs, int n_blocks, int K, int bs, int64_t B,
                           int flags, int precision, pjamd_stream_t stream);

/* Other propagation algorithms of the sum layers (PyJuice's propagation_alg), on the vector ALUs
 * over the same partition tables: alg 1 = "MPE", 2 = "GeneralLL" with exponent alpha > 0, 0 =
 * "LL" (as GeneralLL with alpha = 1, an independent check of the main kernels).  With
 * lt = log(params[pids[j][k] + i]) + element_mars[cids[j][k]][b]:
 * forward: node_mars[nids[j] + i][b] = max_k lt (MPE) or log(sum_k exp(alpha lt)) / alpha.
 * Replaces the propagation_alg_id 1 / 2 branches of SumLayer._fw_triton_block_sparse_*_kernel
 * and _fw_triton_sparse_kernel (sum_layer.py:454-800, 972-1100), selected by
 * Layer.propagation_alg_mapping (layer.py:11-15). */
int pjamd_sum_forward_alg(float* node_mars, const float* element_mars, const float* params, const int* nids,
                          const int* cids, const int* pids, const int* n_chs, int n_blocks, int K, int bs, int64_t B,
                          int alg, float alpha, pjamd_stream_t stream);

/* Element flows for alg 1 / 2 over one backward partition (node flows as stored: these
 * algorithms take no in-place log transform): element_flows[chids[j] + ci][b] = sum over parent
 * rows p of node_flows[p][b] where |lt - node_mars[p][b]| < 1e-6 (MPE), or node_flows[p][b] *
 * exp(alpha (lt - node_mars[p][b])) (GeneralLL), lt = log(theta[p, ci]) + element_mars[chids[j] +
 * ci][b]; accumulate != 0 adds to the stored element flows.  Replaces the propagation_alg_id 1 / 2
 * branches of SumLayer._bk_triton_block_sparse_ele_kernel (sum_layer.py:1569-1730) and the modify
 * pass variants (sum_layer.py:1352-1392). */
int pjamd_sum_backward_ele_alg(float* element_flows, const float* node_flows, const float* node_mars,
                               const float* element_mars, const float* params, const int* chids, const int* parids,
                               const int* parpids, const int* n_pars, int n_cblocks, int KP, int cbs, int bs,
                               int64_t B, int alg, float alpha, int accumulate, pjamd_stream_t stream);

/* Parameter flows for alg 1 / 2 over one forward partition, added atomically:
 * param_flows[pfids[j][k] + i] += sum_b node_flows[n][b] where |lt - node_mars[n][b]| < 1e-6
 * (MPE, not scaled by theta), or theta * sum_b node_flows[n][b] exp(e[b] - node_mars[n][b])
 * (GeneralLL: the LL form, as the reference's kernel computes it); n = nids[j] + i.  pf64 selects
 * a double accumulator; alg | PJAMD_ALG_NEGATE subtracts instead (negate_pflows).  Replaces the
 * propagation_alg_id and negate_pflows branches of SumLayer._bk_triton_block_sparse_par_kernel
 * (sum_layer.py:2049-2169). */
#define PJAMD_ALG_NEGATE 16
int pjamd_sum_backward_par_alg(void* param_flows, int pf64, const float* node_flows, const float* node_mars,
                               const float* element_mars, const float* params, const int* nids, const int* cids,
                               const int* pids, const int* pfids, const int* n_chs, int n_blocks, int K, int bs,
                               int64_t B, int alg, pjamd_stream_t stream);

/* Number of int32 fields per row of the `desc` table below. */
#define PJAMD_PAR_DESC_FIELDS 9

/* pjamd_sum_backward_par for many partitions (of many sum layers) in one launch, after the
 * whole flow backward.  nids / n_chs are the partitions' node-block tables concatenated,
 * cids / pids / pfids their index rows concatenated (each partition keeps its own width K).
 * desc[p][0..8] = {first tile, K, m-tiles per block (bs / tile_rows), child tiles per block,
 * block offset, index-row offset, first element row of the layer, shared-slot flag, copies};
 * tile_desc[t] = p for each of the n_tiles tiles.  copies = c > 1 makes rows p+1 .. p+c-1 tied
 * copies of row p (same K, block count, n_chs, pids and pfids; their own node blocks, child
 * rows and element rows): row p's tiles sum the products of all c copies before writing each
 * slot once, and the copy rows get no tiles of their own.  tile_rows is 128 (block sizes and child
 * counts >= 128, child tiles of 128), 64 (block sizes >= 64, child tiles of 64) or 32 (block size
 * 32, child tiles of 128).  element_mars is the buffer that
 * holds every layer's element rows.  flags as pjamd_sum_backward_par (bit 1 unused; bit 4 applies
 * to the partitions without the shared-slot flag).
 * Same per-partition results as pjamd_sum_backward_par; the reference runs one launch per
 * layer inside SumLayer.backward (sum_layer.py:264-384) under TensorCircuit.backward
 * (tensorcircuit.py:344-538). */
int pjamd_sum_backward_par_batched(void* param_flows, const float* node_flows, const float* node_mars,
                                   const float* element_mars, const float* params, const int* nids, const int* cids,
                                   const int* pids, const int* pfids, const int* n_chs, const int* desc,
                                   const int* tile_desc, int n_tiles, int tile_rows, int64_t B, int flags,
                                   int precision, pjamd_stream_t stream);

/* ----------------------------------------------------------------------------- layer chains */

/* Sizes of the chain tables (int32 fields per operation and per component, int64 fields per
 * pack descriptor) and the limits an operation / component must respect. */
#define PJAMD_CHAIN_OP_INTS 12
#define PJAMD_CHAIN_COMP_INTS 4
#define PJAMD_CHAIN_DESC_INTS 10
#define PJAMD_CHAIN_MAX_OPS 64
#define PJAMD_CHAIN_TAB_INTS 2048
#define PJAMD_CHAIN_MAX_ROWS 256
#define PJAMD_CHAIN_MAX_KC 4
#define PJAMD_CHAIN_COLS 16

/* Packed matrix-core operands of the chain groups (precision 1 = BF16, 2 = FP16), written from
 * the FP32 parameters.  Group i (desc int64 [n][PJAMD_CHAIN_DESC_INTS] =
 * {R, K, log2 bs, log2 cbs, parent blocks, child blocks, forward offset, backward offset,
 * pbase offset, 0}; offsets in 16-byte units) is the dense theta[p][c] = params[pbase[p / bs]
 * [c / cbs] + (c % cbs) bs + p % bs] over R parent rows and K child rows, packed for the
 * 16x16x32 matrix-core instructions: forward unit (t, s, l) = theta[16 t + l % 16]
 * [32 s + 8 (l / 16) + q], backward unit (t, s, l) = theta[32 s + 8 (l / 16) + q][16 t + l % 16],
 * q < 8, units in row-major (t, s, l) order.
 * Workgroup j packs 128 units: group wg[3 j], direction wg[3 j + 1], from unit wg[3 j + 2].
 * No counterpart in the reference, which reads FP32 parameters in every layer kernel
 * (sum_layer.py:454-568); this copy feeds pjamd_chain_forward / pjamd_chain_backward. */
int pjamd_chain_pack(void* packed, const float* params, const int64_t* desc, const int64_t* pbase, const int* wg,
                     int n_wg, int precision, pjamd_stream_t stream);

/* A run of narrow levels (product layer + sum layer above it) in one launch.  The run's groups
 * (dense sum blocks, R <= 256 parent rows over K <= 256 child element rows, K a multiple of 32,
 * block sizes powers of two, bs >= 8, cbs >= 32) are split into n_comp components that never
 * exchange data; workgroup w computes component w % n_comp on batch columns
 * [16 (w / n_comp), 16 (w / n_comp) + 16), its operations in order.  comp int32 [n_comp][4] = {first operation, operations (<= PJAMD_CHAIN_MAX_OPS), first
 * table entry, table entries (<= PJAMD_CHAIN_TAB_INTS)}; ops int32 [n][PJAMD_CHAIN_OP_INTS] =
 * {R, K, log2 bs, log2 cbs, parent blocks, child blocks, KC (<= 4), nid offset, child-block
 * offset, packed operand offset (16-byte units), flags, 0}; table offsets are relative to the
 * component's first entry: nid = first node row of each parent block; child-block entries of
 * 1 + KC ints = {first element row (of the whole element_mars), KC child references: node-row
 * start (0 = none), or -(1 + 256 slot + first position) for rows of one of the component's two
 * previous operations, which the kernel keeps in LDS}; flags bit 3: the operation reads rows its
 * component wrote earlier from global memory.  Forward: element rows = sums of the children's node rows (stored), node rows =
 * log sum theta exp(element rows).  B % 4 == 0.  Replaces the forward loop over the inner layer
 * groups of TensorCircuit.forward (tensorcircuit.py:269-283) for the levels of the run. */
int pjamd_chain_forward(float* node_mars, float* element_mars, const void* packed, const int* comp, const int* ops,
                        const int* tabs, int n_comp, int64_t B, int precision, pjamd_stream_t stream);

/* The backward of a run, top-down, same split and tables; each parent block's node row is
 * followed (after all of them) by its flow source: -1 = node_flows, else 256 slot + first row of
 * the element flows one of the component's two previous operations keeps in LDS; child-block
 * entries hold codes (node row << 2 | marked | 2 when every row of the child is read from LDS by
 * a later operation, which then stores them as log(flow) - mar) instead of node rows.  Per group: element flows of the child rows
 * from the parents' log(flow) - mar (flags bit 0: the parents' flows already hold it; bit 1:
 * they are raw and it is stored back; neither: computed on the fly), written straight into the
 * product children's node flows (every child has one parent), as log(flow) - mar for the
 * marked children when flags bit 2 is set.  Replaces the backward loop over the inner layer
 * groups of TensorCircuit.backward (tensorcircuit.py:436-470) for the levels of the run (their
 * parameter flows come from pjamd_sum_backward_par_batched afterwards). */
int pjamd_chain_backward(float* node_flows, const float* node_mars, const float* element_mars, const void* packed,
                         const int* comp, const int* ops, const int* tabs, int n_comp, int64_t B, int precision,
                         pjamd_stream_t stream);

/* Debug: C[16][16] = A[16][32] B[32][16] (row-major FP32 buffers) through one 16x16x32 BF16
 * (precision 1) or FP16 (2) matrix-core instruction with the lane layout the chain kernels
 * assume.  No counterpart in the reference. */
int pjamd_debug_mfma16_probe(int precision, const float* A, const float* B, float* C, pjamd_stream_t stream);

/* ----------------------------------------------------------------------------- EM */

/* Mini-batch EM of the sum parameters.  recs is int32 [n_recs][5] = {first parameter, first
 * parameter flow, node-block id, block size, number of children} per node block (the node
 * blocks of compile_par_update_fn); each node's parameters become
 * (1 - step) theta + step * (f + pseudocount / n) / (sum f + pseudocount), in double.
 * Replaces em_par_update (src/pyjuice/model/backend/par_update.py:214-257, 296-330). */
int pjamd_sum_em(float* params, const void* pflows, int pf64, const int* recs, int n_recs, float step,
                 float pseudocount, int keep_zero, int max_bs, pjamd_stream_t stream);

/* pflows[target + i] += sum_c pflows[copies[c] + i] for i < n: the flows of tied copies are
 * added to their source.  Replaces compute_cum_par_flows
 * (src/pyjuice/model/backend/parflow_fusing.py:81-120). */
int pjamd_pflow_fuse(void* pflows, int pf64, int64_t target, int64_t n, const int64_t* copies, int ncopies,
                     pjamd_stream_t stream);

/* ----------------------------------------------------------------------------- sampling */

/* Top-down sampling through one sum partition: for each active node (node_act != 0, cleared as
 * it is read) and sample s, select a child and mark element_act[cids[j][k] + ...][s].  The
 * children are scanned in chunks of 512 (the reference's BLOCK_K, sample.py:115-131): within a
 * chunk k is the first child whose FP32 running sum of theta (left to right, from 0) exceeds u;
 * when u reaches the chunk's total, u -= total and the next chunk is scanned; past the last
 * chunk the last real child is taken (where the reference would read index -1).  u is
 * uniforms[node][s] or, when uniforms is NULL, the counter-based uniform of (node, s): word 0 of
 * Philox4x32-10 (Salmon et al., SC 2011) with key (seed[0], seed[1]) and counter (s, node, 0, 0),
 * as (x >> 8) * 2^-24.  seed is a device pointer, so a recorded graph takes a new key per replay.
 * cdf (optional): the running sums of params from pjamd_sample_prefix (per node, chunks of 512,
 * in the order of the table's children), read instead of summing per workgroup; valid when each
 * row's parameters are the run pids[j][0] + k * bs.
 * Replaces sample_sum_layer (src/pyjuice/queries/sample.py:38-181). */
int pjamd_sample_sum(uint8_t* elem_act, uint8_t* node_act, const float* uniforms, const uint32_t* seed,
                     const float* params, const float* cdf, const int* nids, const int* cids, const int* pids,
                     const int* n_chs, int n_blocks, int K, int bs, int64_t S, pjamd_stream_t stream);

/* Running sums of strided rows: out[row_base[r] + c * row_stride[r]] = sum over c' <= c of
 * in[row_base[r] + c' * row_stride[r]], c < row_len[r], in FP32 left to right, restarting at
 * every multiple of chunk (0: never).  With the sum parameters' rows (one per node, stride = block
 * size, chunk 512) it gives the cdf of pjamd_sample_sum; with the Categorical rows (stride 1, no
 * chunks) that of pjamd_input_sample, so sampling kernels do not sum parameters per workgroup.
 * No PyJuice counterpart (its sampling kernels take tl.cumsum per program, sample.py:125). */
int pjamd_sample_prefix(float* out, const float* in, const int64_t* row_base, const int* row_stride,
                        const int* row_len, int64_t n_rows, int chunk, pjamd_stream_t stream);

/* Marks every child of an active product node active and clears the product's flag, so the
 * activity masks are zero again after a pass.  Replaces sample_prod_layer
 * (src/pyjuice/queries/sample.py:270-333). */
int pjamd_sample_prod(uint8_t* node_act, uint8_t* elem_act, const int* nids, const int* cids, int n_blocks,
                      int K, int bs, int64_t S, pjamd_stream_t stream);

/* ----------------------------------------------------------------------------- diagnostics */

/* C = A (32 x KI) * B (KI x 32) on one MFMA instruction (KI = 2 for precision 0, 16 for 1 and 2):
 * checks the operand / accumulator lane layouts the sum kernels assume.  No PyJuice
 * counterpart; used by the tests. */
int pjamd_debug_mfma_probe(int precision, const float* A, const float* B, float* C, pjamd_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* PYJUICE_AMD_H */
