// hf_kernels_inst.h -- launcher definitions, included once per arithmetic
// type by hf_kernels_f64.hip / hf_kernels_f32.hip with HF_MATH set.
#pragma once
#include "hf_kernels.h"
#include "hf_launch.h"

#include <hip/hip_ext.h>

namespace hf {

typedef HF_MATH Math;
typedef Math::R Real;

template <>
hipError_t launch_update<Real>(int scheme, int limiter, bool resid, int kind, const HfStepArgs<Real>& a, int nwg,
                              hipStream_t s, hipEvent_t e0, hipEvent_t e1)
{
    if (a.nitems <= 0) return hipSuccess;
    const dim3 block(256), grid(nwg);
#define HF_GO(K)                                                           \
    do {                                                                   \
        if (e0 || e1) hipExtLaunchKernelGGL((K), grid, block, 0, s, e0, e1, 0, a); \
        else hipLaunchKernelGGL((K), grid, block, 0, s, a);                \
    } while (0)
    if (scheme == 0) {
        if (resid) HF_GO((k_godunov<Math, true>));
        else HF_GO((k_godunov<Math, false>));
    } else if (kind == HF_LAUNCH_GUESS) {
        if (limiter == 0) HF_GO((k_muscl_guess<Math, 0>));
        else HF_GO((k_muscl_guess<Math, 1>));
    } else if (kind == HF_LAUNCH_CHECK) {
        if (limiter == 0) HF_GO((k_muscl_check<Math, 0>));
        else HF_GO((k_muscl_check<Math, 1>));
    } else if (limiter == 0) {
        if (resid) HF_GO((k_muscl<Math, 0, true>));
        else HF_GO((k_muscl<Math, 0, false>));
    } else {
        if (resid) HF_GO((k_muscl<Math, 1, true>));
        else HF_GO((k_muscl<Math, 1, false>));
    }
#undef HF_GO
    return hipGetLastError();
}

namespace {
template <class K>
int occupancy(K kernel)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, 256, 0) != hipSuccess || n < 1) n = 1;
    return n;
}
}  // namespace

template <>
int update_wgs_per_cu<Real>(int scheme, int limiter, bool resid)
{
    if (scheme == 0) return resid ? occupancy(k_godunov<Math, true>) : occupancy(k_godunov<Math, false>);
    if (limiter == 0) return resid ? occupancy(k_muscl<Math, 0, true>) : occupancy(k_muscl<Math, 0, false>);
    return resid ? occupancy(k_muscl<Math, 1, true>) : occupancy(k_muscl<Math, 1, false>);
}

template <>
hipError_t launch_ghost<Real>(const HfGhost* tab, int n, const HfBlockDev* blocks, const Real* S, Real* D,
                              const Real* consts, unsigned long long* gmax, double gamma, const GhostBook& book,
                              hipStream_t s)
{
    const int threads = 64 * GHOST_WAVES;
    const int nb = (n > 0 ? (n + threads - 1) / threads : 0) + 1;   // + the bookkeeping workgroup
    hipLaunchKernelGGL((k_ghost<Math>), dim3(nb), dim3(threads), 0, s, tab, n, blocks, S, D, consts, gmax, gamma,
                       book);
    return hipGetLastError();
}

template <>
hipError_t launch_scan<Real>(const HfBlockDev& bd, const Real* I, const Real* Gh, int interior_only, double gamma,
                             unsigned long long* out, hipStream_t s)
{
    const long long cells = (long long)bd.tx * bd.ty;
    long long nb = (cells + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL((k_scan<Math>), dim3((unsigned)nb), dim3(256), 0, s, bd, I, Gh, interior_only, gamma, out);
    return hipGetLastError();
}

template <>
hipError_t launch_conserve<Real>(const HfBlockDev& bd, const Real* A, double* part, int nwg, hipStream_t s)
{
    hipLaunchKernelGGL((k_conserve<Real>), dim3(nwg), dim3(256), 0, s, bd, A, part);
    return hipGetLastError();
}

}  // namespace hf
