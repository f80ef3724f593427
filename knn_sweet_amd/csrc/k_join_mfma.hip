// k_join_mfma.hip -- launchers of the MFMA join kernels: dispatch on the
// operand width (the kernels themselves are instantiated per width in
// k_join_mfma_d<DK>.hip).
//   PREC 0: plain FP32 (v_mfma_f32_16x16x4_f32)
//   PREC 1: bf16 filter (one v_mfma_f32_16x16x32_bf16 product)
//   PREC 3: compensated bf16 hi/lo filter (three products)
//   PREC 8: fp8 filter (v_mfma_f32_16x16x32_fp8_fp8), widths 64, 96, 128
#include "k_join_mfma.h"

#define KJ_EXTERN_ALL(DKV) KJ_DK_EXTERN(0, DKV) KJ_DK_EXTERN(1, DKV) KJ_DK_EXTERN(3, DKV)
KJ_FOR_WIDTHS(KJ_EXTERN_ALL)
KJ_DK_EXTERN(8, 64)
KJ_DK_EXTERN(8, 96)
KJ_DK_EXTERN(8, 128)

namespace {

template <int PREC>
hipError_t launch_mfma_any(const JoinArgs &a, int64_t grid, hipStream_t s, const JoinArgs *argv = nullptr,
                           const Tile *tiles = nullptr, int32_t *queue = nullptr) {
  if constexpr (PREC == 8) {
    switch (a.dk) {
      case 64: return launch_mfma_dk<8, 64>(a, grid, s, argv, tiles, queue);
      case 96: return launch_mfma_dk<8, 96>(a, grid, s, argv, tiles, queue);
      case 128: return launch_mfma_dk<8, 128>(a, grid, s, argv, tiles, queue);
      default: return hipErrorInvalidValue;
    }
  }
  switch (a.dk) {
#define KJ_CASE(DKV) \
  case DKV: return launch_mfma_dk<PREC, DKV>(a, grid, s, argv, tiles, queue);
    KJ_FOR_WIDTHS(KJ_CASE)
#undef KJ_CASE
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

int join_mfma_split_waves(int kcap, int dk) { return mfma_waves(kcap, dk); }

int join_mfma_tile_q(int kcap, int dk, int want_waves) { return MF_TQ * mfma_launch_waves(kcap, dk, want_waves); }

hipError_t launch_join_mfma_f32(const JoinArgs &a, int64_t max_tiles, hipStream_t s) {
  if (max_tiles <= 0) return hipSuccess;
  return launch_mfma_any<0>(a, max_tiles, s);
}

hipError_t launch_join_mfma_f32_multi(const JoinArgs &proto, const JoinArgs *argv, const Tile *tiles, int nslots,
                                      hipStream_t s, int32_t *queue) {
  if (nslots <= 0) return hipSuccess;
  return launch_mfma_any<0>(proto, nslots, s, argv, tiles, queue);
}

hipError_t launch_join_mfma_fp8(const JoinArgs &a, int64_t max_tiles, hipStream_t s) {
  if (max_tiles <= 0) return hipSuccess;
  return launch_mfma_any<8>(a, max_tiles, s);
}

hipError_t launch_join_mfma_fp8_multi(const JoinArgs &proto, const JoinArgs *argv, const Tile *tiles, int nslots,
                                      hipStream_t s, int32_t *queue) {
  if (nslots <= 0) return hipSuccess;
  return launch_mfma_any<8>(proto, nslots, s, argv, tiles, queue);
}

hipError_t launch_join_mfma_bf16(const JoinArgs &a, int terms, int64_t max_tiles, hipStream_t s) {
  if (max_tiles <= 0) return hipSuccess;
  return terms == 3 ? launch_mfma_any<3>(a, max_tiles, s) : launch_mfma_any<1>(a, max_tiles, s);
}

hipError_t launch_join_mfma_bf16_multi(const JoinArgs &proto, int terms, const JoinArgs *argv, const Tile *tiles,
                                       int nslots, hipStream_t s, int32_t *queue) {
  if (nslots <= 0) return hipSuccess;
  return terms == 3 ? launch_mfma_any<3>(proto, nslots, s, argv, tiles, queue)
                    : launch_mfma_any<1>(proto, nslots, s, argv, tiles, queue);
}
