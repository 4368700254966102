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
    const double* cs;        // [4][cs_terms][n]: cos, sin of the outward normal angle, then (cs_terms 4)
                             // cos, sin of its negative; with cs_terms 2 those are cos and -sin
    const double* inv_vol;   // [n] 1 / cell volume
    int32_t cs_terms, pad2;
    double gamma, cfl;
    double inlet[4];
};

// A slot holds the minimum of dr / (|u| + c) over the cells as the bits of a
// positive double, spread over SLOT_WAYS words on separate 128-byte lines:
// each workgroup reduces its cells and does one atomicMin on word
// (workgroup % SLOT_WAYS); a reader takes the minimum of the words.  Same-
// address atomics from every workgroup serialise at the memory side; spread
// over lines they proceed in parallel.
constexpr int SLOT_WAYS = 64;
constexpr int SLOT_STRIDE = 16;                       // u64 words between ways (128 B)
constexpr int SLOT_WORDS = SLOT_WAYS * SLOT_STRIDE;   // one slot
constexpr unsigned long long SLOT_EMPTY = 0x4341C37937E08000ull;   // bits of 1e16 (timestep.cpp:37)

// One step: U^{n+1} = B from U^n = A.  The dt comes from slot rd unless
// use_fixed; the new state's minimum goes to slot wr (wr < 0: not computed)
// and slot rs is reset for the step after next.
struct UnsStep {
    const double* A;
    double* B;
    unsigned long long* slots;   // [3][SLOT_WORDS]
    int32_t rd, wr, rs, use_fixed;
    double fixed_dt;
    long long kcall;         // TimeStep call index (1-based) of this step, unless kdev is set
    double* hist;
    int32_t hist_cap;
    // graph-captured steps: the call index is kdev[kpar] + 1, and workgroup 0
    // stores it in kdev[kpar ^ 1] for the next step
    int32_t kpar;
    unsigned long long* kdev;
};

// e0/e1 (optional): events stamped at the dispatch's start and end
hipError_t launch_uns_step(const UnsMesh& m, const UnsStep& s, hipStream_t st, hipEvent_t e0 = nullptr,
                           hipEvent_t e1 = nullptr);
// minimum of dr / (|u| + c) of state A into slot `slot` (reset first)
hipError_t launch_uns_scan(const UnsMesh& m, const double* A, unsigned long long* slots, int slot, hipStream_t st);

}  // namespace hfu
