// k_join_mfma_bf16.hip -- BF16 MFMA filter join kernels (v_mfma_f32_16x16x32_bf16).
#include "k_join_mfma.h"

hipError_t launch_join_mfma_bf16(const JoinArgs &a, int64_t max_tiles, hipStream_t s) {
  if (max_tiles <= 0) return hipSuccess;
  return launch_mfma_any<true>(a, max_tiles, s);
}
