// Sum-layer forward on MFMA (BF16 / FP16 operands, FP32 accumulation).
#include "pjamd_common.h"

int pjamd_sum_forward_mfma(float* node_mars, const float* element_mars, const float* params, const int* nids,
                           const int* cids, const int* pids, const int* nch, int n_blocks, int K, int bs, int64_t B,
                           int precision, hipStream_t stream) {
    return PJAMD_EUNSUPPORTED;
}
