// k_join_mfma_d32.hip -- the MFMA join kernels for operand width 32
// (plain FP32, bf16 and bf16 hi/lo filters), one translation unit per width
// so the widths compile in parallel.
#include "k_join_mfma.h"

KJ_DK_INST(0, 32)
KJ_DK_INST(1, 32)
KJ_DK_INST(3, 32)
