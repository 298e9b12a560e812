// Evaluation kernels with contracted FP32 numerics.
#define TG_EXACT 0
#include "tg_eval.inc"
