// hf_uns.h -- structures and launchers of the unstructured quadrilateral-cell
// Godunov solver (HyperFlow's src/fvm_2d_uns) on gfx950.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hfu {

// The mesh on the device.  Per cell-edge arrays are edge-major ([e][n]) so a
// thread per cell reads them coalesced.
struct UnsMesh {
    int32_t n, pad;
    const int32_t* nbr;      // [4][n] neighbour cell or boundary code
    const double* len;       // [4][n] edge length
    const double* cs;        // [4][4][n] cos, sin of the outward normal angle, then of its negative
    const double* inv_vol;   // [n] 1 / cell volume
    const double* dr;        // [n] incircle radius (shortest edge)
    double gamma, cfl;
    double inlet[4];
};

// One step: U^{n+1} = B from U^n = A.  The dt comes from slots[rd] (the
// minimum of dr / (|u| + c) over the cells, as bits of a positive double)
// unless use_fixed; the new state's minimum goes to slots[wr] (wr < 0: not
// computed) and slots[rs] is reset for the step after next.
struct UnsStep {
    const double* A;
    double* B;
    unsigned long long* slots;
    int32_t rd, wr, rs, use_fixed;
    double fixed_dt;
    long long kcall;         // TimeStep call index (1-based) of this step
    double* hist;
    int32_t hist_cap, pad;
};

hipError_t launch_uns_step(const UnsMesh& m, const UnsStep& s, hipStream_t st);
// minimum of dr / (|u| + c) of state A into slots[slot] (reset first)
hipError_t launch_uns_scan(const UnsMesh& m, const double* A, unsigned long long* slots, int slot, hipStream_t st);

}  // namespace hfu
