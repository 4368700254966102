// hf_context.h -- the solver context behind the C ABI, shared by the
// translation units of the library (not installed, not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/hyperflow.h"
#include "hf_common.h"

// Records msg as the calling thread's last error and returns code.
int hf_fail(int code, const std::string& msg);
// True once the process has started exiting (an atexit handler of this
// library, which runs before the HIP runtime's own teardown).  Destroy calls
// made after that release host memory only: the HIP runtime may already be
// shutting down, and the driver reclaims the device memory with the process.
bool hf_process_exiting();

#define HF_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return hf_fail(HF_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

struct hf_ctx {
    int scheme = 0, limiter = 0, prec = 0, dt_mode = 0, g = 1, nb = 0, np = 0;
    double gamma = 1.4, cfl = 0.5;
    size_t esize = 8;
    std::vector<hf_block_desc> bdesc;
    std::vector<HfBlockDev> blocks;
    std::vector<HfProbDev> probs;
    std::vector<HfItem> items;
    std::vector<int> item_begin, item_count;      // per block
    std::vector<std::pair<int, int>> fix_rounds;  // item ranges of the fix-up launches (lb = 1, 2, ..)
    std::vector<int32_t> segs;       // [3][nb][2] (first, count) of each block in the main list's phases
    std::vector<int32_t> fixrange;   // [nb][2] (first, count) of each block in the fix-up rounds
    int main_begin = 0, main_items = 0;           // item range of a step's main launch
    int maxnb = 1;
    int wg_cap = 1024;                            // resident workgroups of an update launch
    unsigned long long wait_ticks = 1000000;      // bound of an in-launch wait (s_memrealtime ticks)
    int probe_max = 0;                            // HfStepArgs::probe_max of the context's launches
    int64_t total = 0;                            // elements per state buffer
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;
    bool ghosts_pending = false;   // buf[cur^1]'s ghosts equal buf[cur]'s (apply_bcs done)
    bool partials_valid = false;   // imax[cur] describes buf[cur], gmax[cur] is zero

    HfBlockDev* d_blocks = nullptr;
    HfProbDev* d_probs = nullptr;
    HfItem* d_items = nullptr;
    HfGhost* d_fill = nullptr;     // ghost-fill table
    HfGhost* d_copy = nullptr;     // ghost identity table (copy ghosts between buffers)
    int nfill = 0, ncopy = 0;
    void* d_consts = nullptr;
    unsigned long long* d_imax = nullptr;   // [2][nb][2]
    unsigned long long* d_gmax = nullptr;   // [2][nb][2]
    unsigned long long* d_scan = nullptr;   // [nb][2] scan results + [nb][2] scratch maxima
    unsigned long long* d_imax_fix = nullptr;   // [2][nb][2] maxima of blocks recomputed by the guess / check launch
    uint32_t* d_fixed = nullptr;                // [2][nb] block recomputed by the guess / check launch
    uint32_t* d_sync = nullptr;                 // [HF_SLOTS][HF_SYNC_WORDS] launch counters (HF_SLOT_MAIN, ..)
    int32_t* d_rounds = nullptr;                // plan_table(): round starts, main-list segments, fix-up ranges
    uint32_t* d_pending = nullptr;              // [npending] bitmap over the fix-up items
    int nrounds = 0, fix_items = 0, npending = 1;
    uint32_t* d_skip = nullptr;                 // [nb] block left to the guess launch this step
    uint32_t* d_skip_next = nullptr;            // [nb]
    unsigned long long* d_fixstats = nullptr;   // [4] fix-up tallies (GhostBook::fixstats)
    uint32_t* d_calls = nullptr;
    uint32_t* h_calls_pinned = nullptr;
    hipEvent_t calls_copied = nullptr;   // the last upload of h_calls_pinned has been read
    double* d_hist = nullptr;
    int hist_cap = 1 << 20;
    int64_t hist_base = 0;          // d_hist[k - 1 - hist_base] holds TimeStep call k of problem 0
    double* d_part = nullptr;
    int part_wg = 256;
    void* d_resid = nullptr;
    int64_t resid_elems = 0;
    std::vector<int64_t> calls;     // host mirror of the TimeStep call counters
    hipStream_t stream = nullptr;
    // captured step loops by (steps, start parity, fixed-dt bits or 0 for the CFL dt)
    std::map<std::tuple<int, int, uint64_t>, std::pair<hipGraph_t, hipGraphExec_t>> graphs;

    int profile = 0;                 // 0 off, else time the launches of every profile-th step
    int64_t prof_step = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending, ev_free;
    int64_t prof_launches = 0;
    double prof_ms = 0.0;

    unsigned long long* imax(int parity) { return d_imax + (size_t)parity * nb * 2; }
    unsigned long long* gmax(int parity) { return d_gmax + (size_t)parity * nb * 2; }
};
