// Evaluation kernels with reference bit parity (built with -ffp-contract=off).
#define TG_EXACT 1
#include "tg_eval.inc"
