// FP32 bulk kernels: flux and update arithmetic in binary32 with hardware
// reciprocal/square root; wave-speed maxima and dt stay FP64.
#define HF_MATH MathF32
#include "hf_kernels_inst.h"
