// hf_cnn.h -- structures and launchers of the learned coarse-grid correction
// (a 5-layer 3x3 CNN on BF16 MFMA with FP32 accumulation).
//
// Activations are NHWC BF16 images with a one-pixel zero border:
// [(ny+2)][(nx+2)][C], C = 8 (the 4 conserved variables, zero padded) or 32.
// The border is zeroed once at allocation and never written, so every 3x3
// tap reads in bounds and "same" zero padding needs no branches.
//
// A convolution is an implicit GEMM per 16-pixel row segment:
//   out[o][p] = sum_k W^T[o][k] * patch[k][p],  k = (tap, input channel),
// run as v_mfma_f32_16x16x32_bf16 with the weights as the A operand (rows =
// output channels, preloaded in registers in the order the lanes consume
// them) and the input patch as the B operand (one 16-byte load of 8
// channels per lane).  The weight gradient is the same product with the
// pixels as the reduction dimension, staged through LDS in channel-major
// order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hfn {

enum { CONV_HIDDEN = 0, CONV_BWD = 1, CONV_OUT = 2, CONV_APPLY = 3 };

struct ConvArgs {
    const uint16_t* in;     // NHWC BF16, padded, CIN channels
    const uint16_t* wa;     // A operands [MT][S][64 lanes][8]
    const float* bias;      // [co_real] (HIDDEN, OUT, APPLY)
    int32_t co_real;
    int32_t nx, ny, rows;   // interior size; rows per wave
    uint16_t* out;          // HIDDEN/BWD: NHWC BF16, padded, COUT channels
    const uint16_t* mask;   // BWD: activation whose positive entries pass the gradient
    float* outf;            // OUT: dense planes [4][ny][nx]
    float* U;               // APPLY: state planes (interior origin), corrected in place
    int64_t plane;          // APPLY: elements per plane of U
    int32_t pitch;          // APPLY: row pitch of U
    float gamma;            // APPLY: ideal-gas gamma for the pressure check
    float scale[4];         // APPLY: correction = scale * network output
    float max_output;       // APPLY: fall back where |output| > max_output (<= 0: no bound)
    const uint8_t* shock;   // APPLY: per-cell shock flag [ny][nx] (null: none)
    uint32_t* fallback;     // APPLY: cells left at the plain state
};

struct WgradArgs {
    const uint16_t* a;      // layer input activation, NHWC padded, CI channels
    const uint16_t* gz;     // gradient at the layer's pre-activation, NHWC padded, CO channels
    float* part;            // [workgroups][9*CI*CO + CO] partial sums
    int32_t nx, ny, rows;
};

// Input normalisation + optional shock sensor.  U: planes of the conserved
// variables with the interior origin at U, element (v, y, x) at
// U[v*plane + y*pitch + x].
struct PackArgs {
    const float* U;
    int64_t plane;
    int32_t pitch, nx, ny;
    float mean[4], iscale[4];
    uint16_t* X;            // NHWC BF16, padded, 8 channels
    uint8_t* shock;         // null: no sensor
    float gamma, theta;
};

struct LossArgs {
    const float* Y;         // network output, dense planes [4][ny][nx]
    const float* Uc;        // plain coarse state, dense planes
    const float* T;         // target state (restricted fine solution), dense planes
    float inv_scale[4];     // target residual = (T - Uc) * inv_scale
    float gscale;           // d loss / d y = gscale * 2 (y - t)
    int32_t nx, ny;
    uint16_t* G;            // NHWC BF16, padded, 8 channels
    float* part;            // [blocks] partial sums of squared errors
};

// Launchers (hf_cnn.hip).  cin/cout are storage channel counts (8 or 32).
hipError_t launch_conv(int cin, int cout, int mode, const ConvArgs& a, hipStream_t s);
hipError_t launch_wgrad(int ci, int co, const WgradArgs& a, int* nwg, hipStream_t s);
int wgrad_workgroups(int nx, int ny, int rows);
hipError_t launch_pack(const PackArgs& a, hipStream_t s);
hipError_t launch_loss(const LossArgs& a, int* nblocks, hipStream_t s);
// tmp: REDUCE_SPLIT * (9*ci*co + co) floats of scratch.
constexpr int REDUCE_SPLIT = 16;
hipError_t launch_reduce_grad(const float* part, int nparts, int ci, int co, int ci_r, int co_r, float* g,
                              float* tmp, hipStream_t s);
hipError_t launch_reduce_sum(const float* part, int n, float scale, float* out, hipStream_t s);
hipError_t launch_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t s);
// Packs the FP32 layer weights W[9][ci_r][co_r] into the A-operand order of
// a convolution with cin -> cout storage channels; transposed = 1 packs the
// data-gradient convolution (taps flipped, channels swapped).
hipError_t launch_pack_weights(int cin, int cout, const float* W, int ci_r, int co_r, int transposed, uint16_t* wa,
                               hipStream_t s);
int packed_weight_elems(int cin, int cout);
hipError_t launch_state_copy(const float* src, int64_t splane, int spitch, float* dst, int64_t dplane, int dpitch,
                             int nx, int ny, hipStream_t s);
hipError_t launch_restrict2(const float* U, int64_t plane, int pitch, int nxc, int nyc, float* dst, hipStream_t s);

}  // namespace hfn
