// hf_launch.h -- host-side launchers of the gfx950 kernels (one translation
// unit per arithmetic type instantiates them).
#pragma once
#include <hip/hip_runtime.h>

#include "hf_common.h"

namespace hf {


// Update launches of a step: the main list, and for the MUSCL compat dt the
// guess and check launches of the fix-up rounds.
enum { HF_LAUNCH_MAIN = 0, HF_LAUNCH_GUESS = 1, HF_LAUNCH_CHECK = 2 };

// e0/e1 (optional): events stamped at the dispatch's start and end by the
// launch itself (hipExtLaunchKernel), for per-launch timing without extra
// commands on the stream.
template <class R>
hipError_t launch_update(int scheme, int limiter, bool resid, int kind, const HfStepArgs<R>& a, int grid,
                         hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// Workgroups (256 threads) of an update kernel that fit on one CU at once.
template <class R>
int update_wgs_per_cu(int scheme, int limiter, bool resid);
template <class R>
hipError_t launch_ghost(const HfGhost* tab, int n, const HfBlockDev* blocks, const R* S, R* D, const R* consts,
                        unsigned long long* gmax, double gamma, const GhostBook& book, hipStream_t s);
template <class R>
hipError_t launch_scan(const HfBlockDev& bd, const R* I, const R* Gh, int interior_only, double gamma,
                       unsigned long long* out, hipStream_t s);
template <class R>
hipError_t launch_conserve(const HfBlockDev& bd, const R* A, double* part, int nwg, hipStream_t s);

}  // namespace hf
