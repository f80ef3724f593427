// k_join_mfma_bf16.hip -- BF16 MFMA filter join kernels (v_mfma_f32_16x16x32_bf16):
// the compensated three-product form (hi/lo operands) and the single product.
#include "k_join_mfma.h"

hipError_t launch_join_mfma_bf16(const JoinArgs &a, int terms, int64_t max_tiles, hipStream_t s) {
  if (max_tiles <= 0) return hipSuccess;
  return terms == 3 ? launch_mfma_any<3>(a, max_tiles, s) : launch_mfma_any<1>(a, max_tiles, s);
}

hipError_t launch_join_mfma_bf16_multi(const JoinArgs &proto, int terms, const JoinArgs *argv, const Tile *tiles,
                                       int nslots, hipStream_t s) {
  if (nslots <= 0) return hipSuccess;
  return terms == 3 ? launch_mfma_any<3>(proto, nslots, s, argv, tiles)
                    : launch_mfma_any<1>(proto, nslots, s, argv, tiles);
}
