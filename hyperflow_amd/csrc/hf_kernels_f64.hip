// FP64 kernels: the reference's arithmetic, bit-exact.  Built with
// -ffp-contract=off so no multiply-add is fused.
#define HF_MATH MathF64
#include "hf_kernels_inst.h"
