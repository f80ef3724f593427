// k_join_mfma_d96.hip -- the MFMA join kernels for operand width 96
// (plain FP32, bf16, bf16 hi/lo and fp8 filters), one translation unit per width
// so the widths compile in parallel.
#include "k_join_mfma.h"

KJ_DK_INST(0, 96)
KJ_DK_INST(1, 96)
KJ_DK_INST(3, 96)
KJ_DK_INST(8, 96)
