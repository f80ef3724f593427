// k_join_mfma_d64.hip -- the MFMA join kernels for operand width 64
// (plain FP32, bf16, bf16 hi/lo and fp8 filters), one translation unit per width
// so the widths compile in parallel.
#include "k_join_mfma.h"

KJ_DK_INST(0, 64)
KJ_DK_INST(1, 64)
KJ_DK_INST(3, 64)
KJ_DK_INST(8, 64)
