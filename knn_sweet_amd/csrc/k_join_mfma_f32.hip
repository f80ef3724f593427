// k_join_mfma_f32.hip -- plain FP32 MFMA join kernels (v_mfma_f32_16x16x4_f32).
#include "k_join_mfma.h"

int join_mfma_split_waves(int kcap, int dk) { return mfma_waves(kcap, dk > 128 ? 256 : 128); }

int join_mfma_tile_q(int kcap, int dk, int want_waves) {
  return MF_TQ * mfma_launch_waves(kcap, dk > 128 ? 256 : 128, want_waves);
}

hipError_t launch_join_mfma_f32(const JoinArgs &a, int64_t max_tiles, hipStream_t s) {
  if (max_tiles <= 0) return hipSuccess;
  return launch_mfma_any<0>(a, max_tiles, s);
}

hipError_t launch_join_mfma_f32_multi(const JoinArgs &proto, const JoinArgs *argv, const Tile *tiles, int nslots,
                                      hipStream_t s) {
  if (nslots <= 0) return hipSuccess;
  return launch_mfma_any<0>(proto, nslots, s, argv, tiles);
}
