/*
 * hyperflow_learned.h -- learned coarse-grid correction for the MI355X
 * block-Cartesian Euler step.
 *
 * No counterpart in the reference: HyperFlow has no learned component.  A
 * small CNN (5 layers of 3x3 convolutions, 4 -> 32 -> 32 -> 32 -> 32 -> 4
 * channels, ReLU between layers; BF16 weights and activations, FP32
 * accumulation and FP32 master weights) corrects the state of a 2x-coarsened
 * MUSCL-Hancock step toward the 2x2 block average of the fine-grid solution:
 *
 *     U <- U + out_scale * net((U - in_mean) * in_scale)
 *
 * cell by cell, except where the corrected density or pressure is not
 * positive, the shock sensor fires on the plain state, or the network output
 * leaves the range it was trained on (there the plain step stands).  The
 * sensor is the normalised second difference of the pressure,
 * |p+ - 2p + p-| / (p+ + 2p + p-), in x or y, above a threshold.
 *
 * Device-pointer arguments (suffix _dev) are dense FP32 planes [4][ny][nx]
 * of (rho, rho*u, rho*v, E).  Network calls run on the stream set with
 * hf_net_set_stream (default: the null stream).  Images need nx a multiple
 * of 64.  Return codes are those of hyperflow.h; hf_last_error() explains.
 */
#ifndef HYPERFLOW_LEARNED_H
#define HYPERFLOW_LEARNED_H

#include "hyperflow.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hf_net hf_net;

typedef struct {
    float in_mean[4];        /* input normalisation */
    float in_scale[4];
    float out_scale[4];      /* correction = out_scale * network output */
    float gamma;             /* ideal-gas gamma for the pressure checks */
    float shock_threshold;   /* sensor threshold; <= 0 disables the sensor */
    float max_output;        /* fall back where |network output| exceeds this; <= 0: no bound */
    float reserved;
    uint64_t seed;           /* He-normal initial weights from this seed */
} hf_net_desc;

HF_API int hf_net_create(const hf_net_desc* desc, hf_net** out);
HF_API void hf_net_destroy(hf_net* net);
HF_API int hf_net_set_stream(hf_net* net, void* hip_stream);
HF_API int hf_net_set_desc(hf_net* net, const hf_net_desc* desc);   /* normalisation, gamma, threshold */
/* Parameters as FP32, layer by layer: W[9][ci][co] (tap = 3*(dy+1)+(dx+1))
 * then b[co].  30084 values. */
HF_API int64_t hf_net_param_count(void);
HF_API int hf_net_set_params(hf_net* net, const float* params, size_t n);
HF_API int hf_net_get_params(hf_net* net, float* params, size_t n);
/* Adam moments and step count back to zero. */
HF_API int hf_net_reset_optimizer(hf_net* net);

/* y_dev = net((u_dev - in_mean) * in_scale), FP32 planes [4][ny][nx]. */
HF_API int hf_net_forward(hf_net* net, const float* u_dev, int nx, int ny, float* y_dev);
/* Gradient of the mean squared error between the network output and
 * (target_dev - u_dev) / out_scale over all cells and the 4 variables, with
 * respect to every parameter (same order as the parameters).  grads may be
 * null; *loss receives the loss (may be null).  Synchronises the stream. */
HF_API int hf_net_gradients(hf_net* net, const float* u_dev, const float* target_dev, int nx, int ny, float* grads,
                            float* loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on that loss.  loss_dev,
 * if not null, receives the loss before the step (device float). */
HF_API int hf_net_train_step(hf_net* net, const float* u_dev, const float* target_dev, int nx, int ny, float lr,
                             float* loss_dev);
HF_API int hf_net_synchronize(hf_net* net);

/* Solver coupling (FP32 contexts).  All run on the context's stream. */
/* Corrects the current state of every block of ctx in place (blocks need nx
 * a multiple of 64); *fallback (may be null; synchronises) receives the
 * number of cells left at the plain state. */
HF_API int hf_correct(hf_ctx* ctx, hf_net* net, int64_t* fallback);
/* nsteps steps with a fixed dt for every problem (both the update and the
 * MUSCL predictor); no TimeStep calls are counted.  HF_RUN_GRAPH captures. */
HF_API int hf_run_fixed(hf_ctx* ctx, int nsteps, double dt, int flags);
/* Interior of a block to / from dense FP32 planes on the device. */
HF_API int hf_state_to_device(hf_ctx* ctx, int block, float* dst_dev);
HF_API int hf_state_from_device(hf_ctx* ctx, int block, const float* src_dev);
/* 2x2 block average of a block's interior (nx, ny even) into dense FP32
 * planes [4][ny/2][nx/2]: the conservative restriction to the coarse grid. */
HF_API int hf_restrict2(hf_ctx* ctx, int block, float* dst_dev);
/* The context's HIP stream (for ordering with other device work). */
HF_API void* hf_stream(hf_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
