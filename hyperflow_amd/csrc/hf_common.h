// hf_common.h -- structures shared by the host library and the gfx950 kernels.
#pragma once
#include <stdint.h>

// One padded block in the state buffers: four planes [var][ty][pitch], x fastest.
struct HfBlockDev {
    int64_t off;      // element offset of plane 0 inside the state buffer
    int64_t plane;    // elements per plane (pitch * ty)
    int32_t nx, ny, tx, ty;
    int32_t pitch;    // row pitch in elements
    int32_t problem;  // owning problem (dt group)
    int32_t lb;       // block index inside its problem
    int32_t g;        // ghost layers
    double dx, dy;
};

// A wave's share of an update launch: one column strip of one block, `rows`
// output rows starting at padded row y0.  Lane l owns padded column x0 + l.
struct HfItem {
    int32_t block, x0, y0, rows;
};

// An independent problem: blocks [first, first+nb) share a dt and a TimeStep
// call counter; cps = TimeStep calls per step (1 Godunov, 1+nb MUSCL).
struct HfProbDev {
    int32_t first, nb, cps, pad;
};

// One ghost cell of the table-driven ghost fill.  The destination gets the
// source cell's value (read from the pre-fill state) or a Dirichlet constant,
// with the x/y momentum optionally negated (composed Reflective sides).
struct HfGhost {
    int32_t dst_block;
    uint32_t dst_cell;   // y * pitch + x in the destination block
    int32_t src_block;   // >= 0: source block; < 0: constant index -src_block-1
    uint32_t src_cell;   // low 30 bits: y * pitch + x; bit 30: negate rho*u; bit 31: negate rho*v
};

#define HF_FLIP_U 0x40000000u
#define HF_FLIP_V 0x80000000u
#define HF_CELL_MASK 0x3fffffffu

// Arguments of one update launch (Godunov or MUSCL-Hancock sweep kernel).
template <class R>
struct HfStepArgs {
    const HfBlockDev* blocks;
    const HfProbDev* probs;
    const HfItem* items;
    int32_t item0, nitems;
    int32_t nprob, nblocks;
    const R* A;   // interior source (U^n)
    const R* G;   // ghost source (filled ghosts)
    R* O;         // destination: U^{n+1} interior, or the residual L
    const unsigned long long* imax_old;  // [nblocks][2] max wave speed of the U^n interior
    unsigned long long* imax_new;        // [nblocks][2] max wave speed of the new interior
    const unsigned long long* gmax;      // [nblocks][2] max wave speed of the ghosts
    const uint32_t* calls;               // TimeStep calls per problem (already bumped for this step)
    double* hist;                        // TimeStep results of problem 0: call k goes to hist[k - 1 - hist_base]
    int64_t hist_base;
    int32_t hist_cap;
    int32_t dt_mode;                     // 0 compat, 1 uniform
    double gamma, cfl;
    int32_t use_fixed;                   // 1: dt_step/dt_pred below instead of the device partials
    int32_t nmain;                       // items [0, nmain) form the main list, the rest the fix-up rounds
    double fixed_step, fixed_pred;
    uint32_t* sync;                      // HF_SYNC_WORDS counters of this launch (zero at launch)
    const int32_t* rounds;               // rounds[l]: first item of fix-up round l (1..nrounds); rounds[nrounds+1] = nitems
    int32_t nrounds, npending;
    // Item ranges (first, count): segs[2 * (phase * nblocks + b)] of block b
    // in each of the main list's three phases, fixrange[2 * b] of block b in
    // the fix-up rounds; nskip: blocks the main list skips this step.
    const int32_t* segs;
    const int32_t* fixrange;
    const uint32_t* nskip;
    unsigned long long wait_ticks;       // bound of an in-launch wait in s_memrealtime ticks (10 ms)
    int32_t probe_max;                   // 1: a wave reads a maxima slot before raising it (raise_max)
    uint32_t* pending;                   // [npending] bitmap over the fix-up items left to the check launch's last wave
    // Fix-up state of the step.  Slot 0 belongs to the guess launch, slot 1 to
    // the check launch: imax_fix[s * 2 * nblocks + 2 * b + axis] holds the
    // maxima of block b recomputed by launch s, fixed[s * nblocks + b] says
    // that it was.
    unsigned long long* imax_fix;
    uint32_t* fixed;
    uint32_t* skip;                      // [nblocks] the main list leaves the block to the guess launch
    uint32_t* skip_next;                 // [nblocks] the exact verdict for the next step
};

// Counters of one update launch: the work-queue ticket counter, the number of
// completed items (bit 31: the abort flag of the bounded waits) and the
// number of waves that have left.
// HF_SYNC_NSKIP (main slot only): blocks the step's main list skips, counted
// by the ghost fill.
enum { HF_SYNC_QUEUE = 0, HF_SYNC_DONE = 1, HF_SYNC_EXITED = 2, HF_SYNC_NSKIP = 3, HF_SYNC_WORDS = 4 };
// Sync slots of a context: the step's main launch, class-level launches, the
// step's guess launch and its check launch.
enum { HF_SLOT_MAIN = 0, HF_SLOT_HOST = 1, HF_SLOT_GUESS = 2, HF_SLOT_CHECK = 3, HF_SLOTS = 4 };
#define HF_SYNC_ABORT_BIT 0x80000000u

// Step bookkeeping done by the first workgroup of a ghost-fill launch.
struct GhostBook {
    uint32_t* calls;
    const HfProbDev* probs;
    int32_t nprob;
    int32_t bump;                 // add cps to every problem's counter
    unsigned long long* clear0;   // arrays to zero (next step's partials)
    unsigned long long* clear1;
    int32_t nclear;               // elements in each
    uint32_t* queue;              // sync counters of this step's update launches (HF_SLOTS x HF_SYNC_WORDS)
    int32_t nqueue;
    uint32_t* pending;            // pending bitmap of this step's check launch
    int32_t npending;             // words
    int32_t nblocks;
    // fold the previous step's fix-ups into its maxima (the check launch's
    // over the guess launch's), then clear them
    unsigned long long* imax_prev;   // maxima of the state this step starts from
    unsigned long long* imax_fix;    // both slots, as in HfStepArgs
    uint32_t* fixed;
    uint32_t* skip;               // skip <- skip_next for the step that starts
    uint32_t* skip_next;
    // optional tallies of the previous step's fix-ups: steps, blocks the main
    // list skipped, blocks the guess launch computed, blocks the check launch
    // computed
    unsigned long long* fixstats;
};
