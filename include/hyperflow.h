/*
 * hyperflow.h -- C ABI of the MI355X (gfx950) block-Cartesian Euler step.
 *
 * This is the drop-in boundary for HyperFlow's src/fvm_2d_cart time step.  The
 * reference has no C ABI: the path sits behind C++ virtual interfaces built by
 * a string-keyed factory.  Each entry point below names the reference
 * interface it replaces (paths relative to the reference checkout).  Per-cell
 * virtual calls cannot cross to a GPU efficiently, so the ABI works at
 * mesh/block granularity.
 *
 * Conventions
 *  - Every int-returning call returns HF_OK (0) or a negative HF_E* code and
 *    never throws; hf_last_error() gives a message for the calling thread.
 *  - Host arrays are caller-owned; device memory is owned by the context.
 *  - Field arrays are conserved variables (rho, rho*u, rho*v, E) of one padded
 *    block, laid out [4][ny+2g][nx+2g] (x fastest), g = 1 for Godunov and 2
 *    for MUSCL-Hancock.  The reference stores the same values as
 *    fields[x][y][var] (src/fvm_2d_cart/mesh/mesh_block/mesh_block.cpp:35).
 *  - One context per host thread; calls on one context are not re-entrant.
 *  - All work of a context runs in order on the context's own HIP stream.
 */
#ifndef HYPERFLOW_H
#define HYPERFLOW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HF_API __attribute__((visibility("default")))

enum { HF_OK = 0, HF_EINVAL = -1, HF_EHIP = -2, HF_ESTATE = -3, HF_ENOMEM = -4 };

/* src/fvm_2d_cart/simulation/parser/parser.cpp:222-242 ("Godunov", "MUSCLHancock") */
enum { HF_GODUNOV = 0, HF_MUSCL_HANCOCK = 1 };
/* parser.cpp:230-233 ("MinBee", "VanLeer") */
enum { HF_MINBEE = 0, HF_VANLEER = 1 };
/* parser.cpp:441-463 ("Transmissive", "Reflective", "Dirichlet", "Internal") */
enum { HF_TRANSMISSIVE = 0, HF_REFLECTIVE = 1, HF_DIRICHLET = 2, HF_INTERNAL = 3 };
/* side order of Mesh::apply_bcs (src/fvm_2d_cart/mesh/mesh/mesh.cpp:33-38) */
enum { HF_BOTTOM = 0, HF_RIGHT = 1, HF_TOP = 2, HF_LEFT = 3 };
/* Arithmetic of the flux/update kernels.  HF_FP64 is the reference's
 * arithmetic (bit-exact); HF_FP32 keeps FP64 for the dt reduction only. */
enum { HF_FP64 = 0, HF_FP32 = 1 };
/* How the MUSCL predictor's dt is obtained inside a step.  HF_DT_COMPAT
 * reproduces the reference (a TimeStep call per block, seeing earlier blocks
 * already updated: muscl_hancock.cpp:246 with simulation.cpp:106-110);
 * HF_DT_UNIFORM uses the step dt for every block.  Both advance the CFL call
 * counter as the reference does.  They coincide for single-block problems. */
enum { HF_DT_COMPAT = 0, HF_DT_UNIFORM = 1 };
/* hf_run flags */
enum { HF_RUN_GRAPH = 1 };

/* One MeshBlock with its four boundary conditions
 * (parser.cpp:262-308 and 432-488). */
typedef struct {
    double x_left, x_right, y_bottom, y_top;
    int32_t nx, ny;          /* interior cells */
    int32_t bc[4];           /* HF_TRANSMISSIVE.. per side, order HF_BOTTOM..HF_LEFT */
    int32_t neighbour[4];    /* block index for HF_INTERNAL sides, else -1 */
    double state[4][4];      /* conserved Dirichlet state per side */
    int32_t problem;         /* independent problem (own dt and CFL counter) */
    int32_t reserved;
} hf_block_desc;

/* A mesh of blocks plus the scheme: the objects the reference's JSONParser
 * factories create (parser.cpp:104-259).  Blocks of one problem must be
 * contiguous; problems are independent simulations advanced together. */
typedef struct {
    int32_t scheme;          /* HF_GODUNOV | HF_MUSCL_HANCOCK */
    int32_t limiter;         /* HF_MINBEE | HF_VANLEER (MUSCL only) */
    int32_t precision;       /* HF_FP64 | HF_FP32 */
    int32_t dt_mode;         /* HF_DT_COMPAT | HF_DT_UNIFORM */
    double gamma;            /* EulerIdealGasModel(gamma), euler_ideal_gas.cpp:14-17 */
    double cfl;              /* TimeStep(model, cfl), timestep.cpp:14-21 */
    int32_t nblocks;
    /* Work plan of the update launches (results do not depend on it; 0 = the
     * library default): rows a wave sweeps per item of the main list, rows per
     * item of the main list's closing phase, rows per item of the fix-up
     * rounds of the compat predictor dt. */
    int32_t tile_rows;
    int32_t tail_rows;
    int32_t fix_rows;
    const hf_block_desc* blocks;
} hf_case_desc;

typedef struct hf_ctx hf_ctx;

/* Builds Mesh + MeshBlocks + BCs + TimeStep + Scheme on the device
 * (replaces JSONParser::generate_mesh/generate_time_step/
 * generate_spatial_scheme, parser.cpp:141-245).  Fields start at zero, as
 * MeshBlock's constructor leaves them (mesh_block.cpp:35). */
HF_API int hf_create(const hf_case_desc* desc, hf_ctx** out);
/* Releases the context and all of its device memory. */
HF_API void hf_destroy(hf_ctx* ctx);
/* MUSCLHancockScheme::ghost_cells / GodunovScheme::ghost_cells
 * (muscl_hancock.cpp:64, godunov.cpp:42). */
HF_API int hf_ghost_cells(const hf_ctx* ctx);

/* MeshBlock::set_fields (mesh_block.cpp:155-158): n = 4*(nx+2g)*(ny+2g). */
HF_API int hf_upload(hf_ctx* ctx, int block, const double* fields, size_t n);
/* MeshBlock::get_fields (mesh_block.cpp:161-164); synchronises the stream. */
HF_API int hf_download(hf_ctx* ctx, int block, double* fields, size_t n);

/* Mesh::apply_bcs (mesh.cpp:31-39): ghost fill of every block, sides in the
 * reference's order with its corner semantics. */
HF_API int hf_apply_bcs(hf_ctx* ctx);
/* TimeStep::operator()(mesh) (timestep.cpp:29-68) for one problem: one
 * counted call over every padded cell of the problem's blocks. */
HF_API int hf_time_step(hf_ctx* ctx, int problem, double* dt);
/* Scheme::operator()(block) (scheme.h:20; godunov.cpp:27-33,
 * muscl_hancock.cpp:41-55): the padded spatial update L of one block, zero in
 * the ghosts.  For MUSCL this makes the predictor's counted TimeStep call. */
HF_API int hf_scheme_residual(hf_ctx* ctx, int block, double* L, size_t n);
/* The per-block loop of Simulation::solve (simulation.cpp:106-110) with
 * ForwardEulerODESolver (forward_euler.cpp:15-33): every block becomes
 * (dt*L)+U, blocks in order.  dt[p] is the step dt of problem p (already
 * clamped by the caller as solve() does).  Uses the ghosts of the last
 * hf_apply_bcs. */
HF_API int hf_advance(hf_ctx* ctx, const double* dt);

/* nsteps iterations of solve()'s step body with the dt kept on the device:
 * apply_bcs, TimeStep, per-block ForwardEuler(Scheme) (simulation.cpp:85-110,
 * without the t_end/frame clamps or output).  Asynchronous: returns once the
 * work is queued on the context stream.  HF_RUN_GRAPH captures the loop in a
 * HIP graph (cached per step count). */
HF_API int hf_run(hf_ctx* ctx, int nsteps, int flags);
/* Waits for the context stream. */
HF_API int hf_synchronize(hf_ctx* ctx);

/* Every TimeStep result of problem 0 so far (step and predictor calls, in
 * call order); *count receives the number of calls made. */
HF_API int hf_dt_history(hf_ctx* ctx, double* out, int max, int* count);
/* Number of TimeStep calls made so far by a problem (TimeStep::steps - 1). */
HF_API int hf_calls(hf_ctx* ctx, int problem, int64_t* calls);

/* Kernel timing with HIP events stamped by the launches themselves: with
 * enable = k > 0 the update-kernel launches of every k-th step of hf_run (and
 * every hf_advance launch) are timed (1 = all; a larger k keeps the timing
 * cost out of most steps); 0 turns timing off.  The stats give the number of
 * timed launches and their summed duration. */
HF_API int hf_profile(hf_ctx* ctx, int enable);
HF_API int hf_kernel_stats(hf_ctx* ctx, int64_t* launches, double* total_ms);
/* Fix-up tallies of the reference-compatible MUSCL dt since the last call
 * (counted by each following step): out[0] steps, out[1] blocks the main
 * launch left to the guess launch, out[2] blocks the guess launch computed,
 * out[3] blocks the check launch recomputed.  No counterpart in the
 * reference: it describes how this implementation meets the reference's
 * sequential block order (simulation.cpp:106-110). */
HF_API int hf_fixup_stats(hf_ctx* ctx, int64_t out[4]);
/* Compensated (Neumaier) FP64 sums of rho, rho*u, rho*v, E over the interior
 * cells of one problem, reduced on the device in a fixed order. */
HF_API int hf_conservation(hf_ctx* ctx, int problem, double sums[4]);

/* One step of hf_run (apply_bcs, TimeStep, the per-block ForwardEuler(Scheme)
 * loop) without a context and without allocating device memory.  No
 * counterpart in the reference: this is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).
 *  - d_in / d_out: device arrays of the blocks back to back, each block
 *    [4][ny+2g][nx+2g] (x fastest), double for HF_FP64 and float for HF_FP32.
 *    d_in is only read; d_out receives the state hf_run leaves after one
 *    step (new interior, ghosts filled from d_in).  Alternate them to step on.
 *  - calls: host array, one TimeStep call counter per problem (the calls
 *    made so far; 0 for a fresh start), advanced by this step's calls.
 *  - d_dt: NULL or a device array receiving problem 0's TimeStep results of
 *    this step in call order (1 for Godunov, 1 + blocks for MUSCL-Hancock).
 *  - stream: a hipStream_t, or NULL for the null stream.  Returns when the
 *    step is complete.
 * Tables and working arrays live in one static device array per GPU (32 MiB);
 * a case that needs more is rejected with HF_EINVAL.  Every call on a GPU
 * shares that array: calls that overlap on one GPU (other threads or
 * streams) race and must be serialised by the caller. */
HF_API int hf_step_noalloc(const hf_case_desc* desc, const void* d_in, void* d_out, int64_t* calls, double* d_dt,
                           void* stream);

/* Message of the last failed call on this thread ("" if none). */
HF_API const char* hf_last_error(void);
/* Library/device description, e.g. "hyperflow_amd gfx950". */
HF_API const char* hf_version(void);

#ifdef __cplusplus
}
#endif
#endif
