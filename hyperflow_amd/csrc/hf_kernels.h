// hf_kernels.h -- gfx950 kernels of the block-Cartesian Euler step.
//
// Update kernels ("sweeps"): each 64-lane wave owns a strip of padded columns
// (lane l <-> column x0+l) and walks down `rows` output rows of one block.
// The reconstruction window, face states, both HLLC fluxes and the update of
// a row live in registers; x neighbours come from cross-lane shuffles, y
// neighbours from the previous iteration.  Every conserved value is read once
// from HBM (plus the 2-lane halo at each strip edge and 2 halo rows per
// strip) and every output written once; nothing intermediate goes to memory.
// The new state's max wave speeds (for the next dt) are reduced in the same
// pass: wave shuffles, then one 64-bit atomic max per wave.
//
// The interior is read from A and the filled ghosts from G (the ghost fill
// writes them into the destination buffer), so no launch ever reads a cell
// that another wave of the same step writes.
#pragma once
#include <hip/hip_runtime.h>

#include "hf_common.h"
#include "hf_physics.h"

namespace hf {

template <class R>
__device__ __forceinline__ void load_cell(const R* __restrict__ base, int64_t plane, R u[4])
{
    u[0] = base[0];
    u[1] = base[plane];
    u[2] = base[2 * plane];
    u[3] = base[3 * plane];
}

// Whole-wave lane shifts by one lane through DPP (wave_shr:1 / wave_shl:1):
// lane l receives lane l-1 (up) or lane l+1 (down); the end lanes keep their
// own value.  No LDS round trip, unlike a bpermute shuffle.
__device__ __forceinline__ int dpp_up(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_dn(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ double shfl_up1(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = dpp_up((int)(b & 0xffffffffll)), hi = dpp_up((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double shfl_dn1(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = dpp_dn((int)(b & 0xffffffffll)), hi = dpp_dn((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float shfl_up1(float x) { return __int_as_float(dpp_up(__float_as_int(x))); }
__device__ __forceinline__ float shfl_dn1(float x) { return __int_as_float(dpp_dn(__float_as_int(x))); }

// Work items of a persistent launch: wave w first takes item w for w below
// nstatic (no atomic, so a launch does not open with every wave hitting one
// counter), then tickets from the queue counter, offset by nstatic.  Every
// lane executes the atomic (only lane 0 adds 1) so that no lane-divergent
// region surrounds the loop's exit test; lane 0's old value is the wave's
// ticket.  A lane-0-only branch here let the compiler's control-flow
// structurizer keep lanes 1..63 in the loop after lane 0 left it.
__device__ __forceinline__ unsigned total_waves() { return gridDim.x * (blockDim.x >> 6); }
__device__ __forceinline__ unsigned wave_id()
{
    // wave-uniform, but derived from threadIdx: say so, or the whole item loop
    // is compiled as divergent (vector loads of the item and block records)
    return __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
}
__device__ __forceinline__ unsigned next_ticket(uint32_t* sync, int lane, unsigned nstatic)
{
    const unsigned t = atomicAdd(&sync[HF_SYNC_QUEUE], lane == 0 ? 1u : 0u);
    return __builtin_amdgcn_readfirstlane(t) + nstatic;
}
__device__ __forceinline__ unsigned first_ticket(uint32_t* sync, int lane, unsigned nstatic)
{
    const unsigned w = wave_id();
    return w < nstatic ? w : next_ticket(sync, lane, nstatic);
}

// One item of the launch is complete: its stores and maxima atomics are made
// visible at device scope before the count that waiting waves poll.
__device__ __forceinline__ void item_done(uint32_t* sync, int lane, unsigned n = 1)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(&sync[HF_SYNC_DONE], lane == 0 ? n : 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded wait until `target` items of the launch are complete (every item
// before position `target` of the queue: items are handed out in order and an
// item that waits never completes before the items it waits for).  The wait
// gives up after WAIT_TICKS of the 100 MHz real-time counter, or at once when
// another wave has given up, and then returns false with the abort bit set:
// the caller leaves its item to the launch's last wave.  One load per poll
// (count and abort bit share a word), about 0.85 us apart.
constexpr unsigned long long WAIT_TICKS = 1000000ull;   // 10 ms
__device__ __forceinline__ bool wait_done(uint32_t* sync, unsigned target)
{
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    for (;;) {
        const unsigned d = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(&sync[HF_SYNC_DONE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (d & HF_SYNC_ABORT_BIT) return false;
        if (d >= target) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            return true;
        }
        if (__builtin_amdgcn_s_memrealtime() - t0 > WAIT_TICKS) {
            __hip_atomic_fetch_or(&sync[HF_SYNC_DONE], HF_SYNC_ABORT_BIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(32);
    }
}

// A wave-reduced value as a wave-uniform (scalar) bit pattern.
__device__ __forceinline__ unsigned long long uniform_bits(double x)
{
    const unsigned long long b = dbits(x);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// Per-wave maxima into the block's slots; every lane holds the reduced value,
// so every lane may issue the (idempotent) atomic -- no divergent region.  The
// value is passed as a scalar: the compiler then issues one atomic from one
// lane, instead of first reducing 64 lane values it cannot tell are equal (a
// 64-iteration scalar loop per atomic).
__device__ __forceinline__ void store_maxima(unsigned long long* out, int block, double mx, double my)
{
    atomicMax(&out[2 * block], uniform_bits(mx));
    atomicMax(&out[2 * block + 1], uniform_bits(my));
}

// Load row y of this lane's column: ghost rows/columns come from the ghost
// source, interior cells from the interior source.
template <class R>
__device__ __forceinline__ void load_row(const R* __restrict__ Acol, const R* __restrict__ Gcol, bool gcol, int y,
                                         int g, int ty, int pitch, int64_t plane, R u[4])
{
    const bool gh = gcol || y < g || y >= ty - g;
    load_cell<R>((gh ? Gcol : Acol) + (int64_t)y * pitch, plane, u);
}

// ---------------------------------------------------------- MUSCL-Hancock --
// MUSCLHancockScheme::operator() (scheme/muscl_hancock/muscl_hancock.cpp:41-55)
// fused with ForwardEulerODESolver (scheme/ode/forward_euler/forward_euler.cpp:25-30).
// Output columns per wave: 60 (lanes 2..61); lanes 0,1,62,63 are halo.
template <class M, int LIM, bool RESID>
__device__ __forceinline__ void muscl_item(const HfStepArgs<typename M::R>& a, const HfItem it, const int lane,
                                           const double dt_step, const double dt_pred,
                                           unsigned long long* __restrict__ imax_out)
{
    typedef typename M::R R;

    const HfBlockDev bd = a.blocks[it.block];
    const int xraw = it.x0 + lane;
    const bool valid = xraw < bd.tx;
    const int xc = valid ? xraw : bd.tx - 1;
    const bool gcol = xc < 2 || xc >= bd.tx - 2;
    const R* Acol = a.A + bd.off + xc;
    const R* Gcol = a.G + bd.off + xc;
    const int pitch = bd.pitch, ty = bd.ty;
    const int64_t plane = bd.plane;
#define ROW(y, u) load_row<R>(Acol, Gcol, gcol, (y), 2, ty, pitch, plane, (u))
    const bool out = valid && lane >= 2 && lane < 62 && xraw >= 2 && xraw < bd.tx - 2;

    const R hx = R((0.5 * dt_pred) / bd.dx);     // half_dt_dx, muscl_hancock.cpp:248
    const R hy = R((0.5 * dt_pred) / bd.dy);
    const R rdx = R(1.0 / bd.dx);                  // one_dx, muscl_hancock.cpp:352
    const R rdy = R(1.0 / bd.dy);
    const R dtR = R(dt_step);
    const R gam = R(a.gamma), g1 = R(a.gamma - 1.0);

    const int y0 = it.y0;
    R um[4], uc[4], up[4], un[4];
    ROW(y0 - 2, um);
    ROW(y0 - 1, uc);
    ROW(y0, up);
    ROW(y0 + 1, un);

    // warm-up: T' of cell y0-1, (B', T') of cell y0, flux through face y0-1/2
    R Gs[4], Tc[4];
    {
        R lo[4], hi[4];
        reconstruct<M, LIM, false>(um, uc, up, hy, g1, lo, hi);   // cell y0-1
        R lo2[4];
        reconstruct<M, LIM, false>(uc, up, un, hy, g1, lo2, Tc);  // cell y0
        const State<R> sl = make_state<M>(hi, gam, g1);
        const State<R> sr = make_state<M>(lo2, gam, g1);
        hllc<M, false>(sl, sr, Gs);
    }
    // window: uc = U(y), up = U(y+1), un = U(y+2) (loaded one iteration ahead
    // so that HBM latency overlaps the row's arithmetic)
    for (int k = 0; k < 4; k++) { uc[k] = up[k]; up[k] = un[k]; }
    ROW(y0 + 2, un);

    R* Ob = RESID ? a.O : a.O + bd.off;
    double mx = 0.0, my = 0.0;
    for (int r = 0; r < it.rows; r++) {
        const int y = y0 + r;
        R nn[4];
        ROW(y + 3 < ty ? y + 3 : ty - 1, nn);     // next iteration's U(y+2)

        // y direction: reconstruct cell y+1, flux through face y+1/2
        R Bn[4], Tn[4], Gn[4];
        reconstruct<M, LIM, false>(uc, up, un, hy, g1, Bn, Tn);
        {
            const State<R> sl = make_state<M>(Tc, gam, g1);
            const State<R> sr = make_state<M>(Bn, gam, g1);
            hllc<M, false>(sl, sr, Gn);
        }

        // x direction: reconstruct cell (x, y) from lane neighbours
        R ul[4], ur[4], Lf[4], Rf[4], Rl[4], Fw[4], Fe[4];
        for (int k = 0; k < 4; k++) { ul[k] = shfl_up1(uc[k]); ur[k] = shfl_dn1(uc[k]); }
        reconstruct<M, LIM, true>(ul, uc, ur, hx, g1, Lf, Rf);
        for (int k = 0; k < 4; k++) Rl[k] = shfl_up1(Rf[k]);
        {
            const State<R> sl = make_state<M>(Rl, gam, g1);
            const State<R> sr = make_state<M>(Lf, gam, g1);
            hllc<M, true>(sl, sr, Fw);          // face x-1/2
        }
        for (int k = 0; k < 4; k++) Fe[k] = shfl_dn1(Fw[k]);   // face x+1/2

        // calculate_spatial_update (muscl_hancock.cpp:364-371) and (dt*L)+U
        R Ln[4], Un[4];
        update_cell<M>(rdx, rdy, dtR, Fw, Fe, Gs, Gn, uc, Ln, Un);
        if (out) {
            R* o = Ob + (int64_t)y * pitch + xraw;
            for (int k = 0; k < 4; k++) o[k * plane] = RESID ? Ln[k] : Un[k];
            if (!RESID) {
                double dx_, dy_;
                wave_speeds<M>(Un[0], Un[1], Un[2], Un[3], gam, g1, dx_, dy_);
                mx = dmax_nan(mx, dx_);
                my = dmax_nan(my, dy_);
            }
        }
        for (int k = 0; k < 4; k++) { Gs[k] = Gn[k]; Tc[k] = Tn[k]; uc[k] = up[k]; up[k] = un[k]; un[k] = nn[k]; }
    }
    if (!RESID) store_maxima(imax_out, it.block, wave_max(mx), wave_max(my));
#undef ROW
}

// The same update for a narrow column band at a block's right edge, with the
// roles of x and y exchanged: lane l owns padded row it.y0 + l and the wave
// walks along the band's columns.  A block whose width leaves a last column
// strip with only a few output columns (1024 = 17 x 60 + 4) would otherwise
// spend a whole wave per row on them.  Every face state, flux and update is
// the same function of the same operands as in muscl_item, in the same order,
// so the results are identical.  Item encoding: rows = -(cols + 256 * nrows),
// x0 = first output column (padded), y0 = padded row of lane 0; lanes 2..61
// write, up to nrows of them.
__device__ __forceinline__ bool item_transposed(const HfItem& it) { return it.rows < 0; }

template <class M, int LIM, bool RESID>
__device__ __forceinline__ void muscl_item_t(const HfStepArgs<typename M::R>& a, const HfItem it, const int lane,
                                             const double dt_step, const double dt_pred,
                                             unsigned long long* __restrict__ imax_out)
{
    typedef typename M::R R;

    const HfBlockDev bd = a.blocks[it.block];
    const int code = -it.rows;
    const int cols = code & 255, nrows = code >> 8;
    const int yraw = it.y0 + lane;
    const bool valid = yraw < bd.ty;
    const int yc = valid ? yraw : bd.ty - 1;
    const bool grow = yc < 2 || yc >= bd.ty - 2;
    const int pitch = bd.pitch, tx = bd.tx;
    const R* Arow = a.A + bd.off + (int64_t)yc * pitch;
    const R* Grow = a.G + bd.off + (int64_t)yc * pitch;
    const int64_t plane = bd.plane;
#define COL(x, u) load_cell<R>(((grow || (x) < 2 || (x) >= tx - 2) ? Grow : Arow) + (x), plane, (u))
    const bool out = valid && lane >= 2 && lane < 62 && lane < 2 + nrows && yraw >= 2 && yraw < bd.ty - 2;

    const R hx = R((0.5 * dt_pred) / bd.dx);     // half_dt_dx, muscl_hancock.cpp:248
    const R hy = R((0.5 * dt_pred) / bd.dy);
    const R rdx = R(1.0 / bd.dx);                  // one_dx, muscl_hancock.cpp:352
    const R rdy = R(1.0 / bd.dy);
    const R dtR = R(dt_step);
    const R gam = R(a.gamma), g1 = R(a.gamma - 1.0);

    const int x0 = it.x0;
    R um[4], uc[4], up[4], un[4];
    COL(x0 - 2, um);
    COL(x0 - 1, uc);
    COL(x0, up);
    COL(x0 + 1, un);

    // warm-up: R' of column x0-1, (L', R') of column x0, flux through face x0-1/2
    R Fw[4], Rc[4];
    {
        R lo[4], hi[4];
        reconstruct<M, LIM, true>(um, uc, up, hx, g1, lo, hi);    // column x0-1
        R lo2[4];
        reconstruct<M, LIM, true>(uc, up, un, hx, g1, lo2, Rc);   // column x0
        const State<R> sl = make_state<M>(hi, gam, g1);
        const State<R> sr = make_state<M>(lo2, gam, g1);
        hllc<M, true>(sl, sr, Fw);
    }
    for (int k = 0; k < 4; k++) { uc[k] = up[k]; up[k] = un[k]; }
    COL(x0 + 2, un);

    R* Ob = RESID ? a.O : a.O + bd.off;
    double mx = 0.0, my = 0.0;
    for (int c = 0; c < cols; c++) {
        const int x = x0 + c;
        R nn[4];
        COL(x + 3 < tx ? x + 3 : tx - 1, nn);     // next iteration's U(x+2)

        // x direction: reconstruct column x+1, flux through face x+1/2
        R Ln[4], Rn[4], Fe[4];
        reconstruct<M, LIM, true>(uc, up, un, hx, g1, Ln, Rn);
        {
            const State<R> sl = make_state<M>(Rc, gam, g1);
            const State<R> sr = make_state<M>(Ln, gam, g1);
            hllc<M, true>(sl, sr, Fe);
        }

        // y direction: reconstruct cell (x, y) from lane neighbours
        R ub[4], ut[4], Bf[4], Tf[4], Tl[4], Gs[4], Gn[4];
        for (int k = 0; k < 4; k++) { ub[k] = shfl_up1(uc[k]); ut[k] = shfl_dn1(uc[k]); }
        reconstruct<M, LIM, false>(ub, uc, ut, hy, g1, Bf, Tf);
        for (int k = 0; k < 4; k++) Tl[k] = shfl_up1(Tf[k]);
        {
            const State<R> sl = make_state<M>(Tl, gam, g1);
            const State<R> sr = make_state<M>(Bf, gam, g1);
            hllc<M, false>(sl, sr, Gs);         // face y-1/2
        }
        for (int k = 0; k < 4; k++) Gn[k] = shfl_dn1(Gs[k]);   // face y+1/2

        // calculate_spatial_update (muscl_hancock.cpp:364-371) and (dt*L)+U
        R Lu[4], Un[4];
        update_cell<M>(rdx, rdy, dtR, Fw, Fe, Gs, Gn, uc, Lu, Un);
        if (out) {
            R* o = Ob + (int64_t)yraw * pitch + x;
            for (int k = 0; k < 4; k++) o[k * plane] = RESID ? Lu[k] : Un[k];
            if (!RESID) {
                double dx_, dy_;
                wave_speeds<M>(Un[0], Un[1], Un[2], Un[3], gam, g1, dx_, dy_);
                mx = dmax_nan(mx, dx_);
                my = dmax_nan(my, dy_);
            }
        }
        for (int k = 0; k < 4; k++) { Fw[k] = Fe[k]; Rc[k] = Rn[k]; uc[k] = up[k]; up[k] = un[k]; un[k] = nn[k]; }
    }
    if (!RESID) store_maxima(imax_out, it.block, wave_max(mx), wave_max(my));
#undef COL
}

// The fix-up verdict of the lb-th block of problem p: the step dt, the
// speculative predictor dt the main list used (every block at U^n) and the
// reference's (blocks before this one at their final U^{n+1}); the block must
// be (re)computed in its fix-up round if the main list skipped it or the two
// predictor dts differ.  Also records the verdict for the next step.
template <class R>
__device__ __forceinline__ bool fix_verdict(const HfStepArgs<R>& a, int p, const HfProbDev& pr, int lb, double& dt_step,
                                            double& dt_pred, long long& kstep, long long& kpred)
{
    double spec;
    step_dts_fix<R>(a, p, pr, lb, dt_step, spec, dt_pred, kstep, kpred);
    const int b = pr.first + lb;
    a.skip_next[b] = spec != dt_pred ? 1u : 0u;   // the same value from every wave
    return a.skip[b] != 0u || spec != dt_pred;
}

// Fix-up rounds from round l on, while they are empty: round l holds the
// lb = l-th block of every problem with more than l blocks, and its verdicts
// need the blocks before it final, which holds while every earlier round was
// empty.  The queue counter jumps past each empty round (its unissued items
// are counted complete).  Returns the first round with work, or nrounds + 1.
template <class R>
__device__ int first_busy_round(const HfStepArgs<R>& a, int l, int lane, unsigned nstatic)
{
    for (; l <= a.nrounds; l++) {
        bool need = false;
        for (int p = 0; p < a.nprob; p++) {
            const HfProbDev pr = a.probs[p];
            if (pr.nb <= l) continue;
            double st, pd;
            long long ks, kp;
            need = fix_verdict<R>(a, p, pr, l, st, pd, ks, kp) || need;
        }
        if (need) return l;
        const unsigned end = (unsigned)a.rounds[l + 1] - nstatic;
        const unsigned old = __builtin_amdgcn_readfirstlane(atomicMax(&a.sync[HF_SYNC_QUEUE], end));
        if (old < end) item_done(a.sync, lane, end - old);
    }
    return a.nrounds + 1;
}

// Which fix-up round an item position belongs to (rounds are few).
template <class R>
__device__ __forceinline__ int round_of(const HfStepArgs<R>& a, unsigned idx)
{
    int l = 1;
    while (l < a.nrounds && idx >= (unsigned)a.rounds[l + 1]) l++;
    return l;
}

// Persistent MUSCL-Hancock launch: waves take work items from a queue until
// it is empty.  The queue holds the main list (every block with the
// speculative predictor dt, all blocks at U^n, except blocks marked skip[b]:
// their speculation failed in the previous step), then the fix-up rounds of
// the reference-compatible dt (muscl_hancock.cpp:246 with the sequential block
// loop of simulation.cpp:106-110): round l recomputes the l-th block of each
// problem with the reference's predictor dt, earlier blocks at their final
// U^{n+1}, wherever the main list skipped it or its speculative dt differs.
// An item of round l first waits (bounded) until every earlier item of the
// queue is complete.  Empty rounds are skipped as a whole.  A wait that gives
// up leaves its item to the launch's last wave, which computes every such
// item in queue order once all other waves have left: no wait there, so the
// launch always ends, and the results are the same bit for bit.
template <class M, int LIM, bool RESID>
__device__ __forceinline__ void muscl_go(const HfStepArgs<typename M::R>& a, const HfItem& it, int lane, double dt_step,
                                         double dt_pred, unsigned long long* imax_out)
{
    if (!RESID && item_transposed(it))
        muscl_item_t<M, LIM, RESID>(a, it, lane, dt_step, dt_pred, imax_out);
    else
        muscl_item<M, LIM, RESID>(a, it, lane, dt_step, dt_pred, imax_out);
}

template <class R>
__device__ __forceinline__ void record_dts(const HfStepArgs<R>& a, const HfItem& it, const HfBlockDev& bd, bool step,
                                           double dt_step, double dt_pred, long long kstep, long long kpred)
{
    if (bd.problem == 0 && it.x0 == 0 && it.y0 == bd.g) {   // same value from every lane
        const long long hs = kstep - a.hist_base, hp = kpred - a.hist_base;
        if (step && bd.lb == 0 && hs >= 1 && hs <= a.hist_cap) a.hist[hs - 1] = dt_step;
        if (hp >= 1 && hp <= a.hist_cap) a.hist[hp - 1] = dt_pred;
    }
}

template <class M, int LIM, bool RESID>
__global__ __launch_bounds__(256) void k_muscl(const HfStepArgs<typename M::R> a)
{
    typedef typename M::R R;
    const int lane = threadIdx.x & 63;
    const unsigned nstatic = min(total_waves(), (unsigned)a.nmain);
    // the block whose dts this wave has computed: speculative ones for the main
    // list (c_fix false) or the fix-up verdict and exact ones (c_fix true)
    int cached = -1;
    bool c_fix = false, c_need = false;
    double c_step = 0.0, c_pred = 0.0;
    long long c_kstep = 0, c_kpred = 0;
    int r_from = -1, r_busy = 0;   // first_busy_round(r_from) as this wave last computed it
    for (unsigned idx = first_ticket(a.sync, lane, nstatic);; idx = next_ticket(a.sync, lane, nstatic)) {
        if (idx >= (unsigned)a.nitems) break;
        const HfItem it = a.items[a.item0 + idx];
        if (a.use_fixed) {
            muscl_go<M, LIM, RESID>(a, it, lane, a.fixed_step, a.fixed_pred, a.imax_new);
            item_done(a.sync, lane);
            continue;
        }
        const HfBlockDev bd = a.blocks[it.block];
        if (idx < (unsigned)a.nmain) {
            if (!a.skip[it.block]) {   // a skipped block is computed exactly in its fix-up round
                if (it.block != cached || c_fix) {
                    cached = it.block;
                    c_fix = false;
                    step_dts<R>(a, it.block, bd, true, DT_SPEC, c_step, c_pred, c_kstep, c_kpred);
                }
                record_dts<R>(a, it, bd, true, c_step, c_pred, c_kstep, c_kpred);
                muscl_go<M, LIM, RESID>(a, it, lane, c_step, c_pred, a.imax_new);
            }
            item_done(a.sync, lane);
            continue;
        }
        // fix-up round item
        const int l = round_of<R>(a, idx);
        if (l != r_from) {
            if (!wait_done(a.sync, (unsigned)a.rounds[l])) {
                atomicOr(&a.pending[(idx - a.nmain) >> 5], 1u << ((idx - a.nmain) & 31));
                continue;   // left to the last wave, not counted complete
            }
            r_from = l;
            r_busy = first_busy_round<R>(a, l, lane, nstatic);
        }
        if (r_busy == l) {
            if (it.block != cached || !c_fix) {
                cached = it.block;
                c_fix = true;
                const HfProbDev pr = a.probs[bd.problem];
                c_need = fix_verdict<R>(a, bd.problem, pr, bd.lb, c_step, c_pred, c_kstep, c_kpred);
            }
            if (c_need) {
                a.fixed[it.block] = 1u;
                record_dts<R>(a, it, bd, false, c_step, c_pred, c_kstep, c_kpred);
                muscl_go<M, LIM, RESID>(a, it, lane, c_step, c_pred, a.imax_fix);
            }
        }
        item_done(a.sync, lane);
    }
    // The last wave to leave computes the items whose wait gave up, in queue
    // order; every other wave has left, so all it reads is final.
    const unsigned left = __builtin_amdgcn_readfirstlane(atomicAdd(&a.sync[HF_SYNC_EXITED], lane == 0 ? 1u : 0u));
    if (left != total_waves() - 1) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (!(__builtin_amdgcn_readfirstlane(a.sync[HF_SYNC_DONE]) & HF_SYNC_ABORT_BIT)) return;
    const unsigned nfix = (unsigned)(a.nitems - a.nmain);
    for (unsigned w = 0; w * 32 < nfix; w++) {
        const unsigned bits = __builtin_amdgcn_readfirstlane(a.pending[w]);
        for (unsigned k = 0; k < 32; k++) {
            if (!((bits >> k) & 1u)) continue;
            const unsigned idx = a.nmain + w * 32 + k;
            const HfItem it = a.items[a.item0 + idx];
            const HfBlockDev bd = a.blocks[it.block];
            const HfProbDev pr = a.probs[bd.problem];
            double dt_step, dt_pred;
            long long kstep, kpred;
            if (fix_verdict<R>(a, bd.problem, pr, bd.lb, dt_step, dt_pred, kstep, kpred)) {
                a.fixed[it.block] = 1u;
                record_dts<R>(a, it, bd, false, dt_step, dt_pred, kstep, kpred);
                muscl_go<M, LIM, RESID>(a, it, lane, dt_step, dt_pred, a.imax_fix);
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");   // its maxima before the next verdict
            }
        }
        a.pending[w] = 0u;
    }
}

// ---------------------------------------------------------------- Godunov --
// GodunovScheme::operator() (scheme/godunov/godunov.cpp:27-118) fused with
// the forward Euler update.  Each cell's primitives and sound speed are
// computed once and shared by its four faces.  Output columns per wave: 62.
template <class M, bool RESID>
__device__ __forceinline__ void godunov_item(const HfStepArgs<typename M::R>& a, const HfItem it, const int lane,
                                             const double dt_step, unsigned long long* __restrict__ imax_out)
{
    typedef typename M::R R;

    const HfBlockDev bd = a.blocks[it.block];
    const int xraw = it.x0 + lane;
    const bool valid = xraw < bd.tx;
    const int xc = valid ? xraw : bd.tx - 1;
    const bool gcol = xc < 1 || xc >= bd.tx - 1;
    const R* Acol = a.A + bd.off + xc;
    const R* Gcol = a.G + bd.off + xc;
    const int pitch = bd.pitch, ty = bd.ty;
    const int64_t plane = bd.plane;
#define ROW(y, u) load_row<R>(Acol, Gcol, gcol, (y), 1, ty, pitch, plane, (u))
    const bool out = valid && lane >= 1 && lane < 63 && xraw >= 1 && xraw < bd.tx - 1;

    const R rdx = R(1.0 / bd.dx);                  // godunov.cpp:97-98
    const R rdy = R(1.0 / bd.dy);
    const R dtR = R(dt_step);
    const R gam = R(a.gamma), g1 = R(a.gamma - 1.0);

    const int y0 = it.y0;
    R ub[4], uc[4];
    ROW(y0 - 1, ub);
    ROW(y0, uc);
    State<R> sb = make_state<M>(ub, gam, g1);
    State<R> sc = make_state<M>(uc, gam, g1);
    R Gs[4];
    hllc<M, false>(sb, sc, Gs);                    // face y0-1/2

    R un[4];
    ROW(y0 + 1, un);                               // loaded one iteration ahead
    R* Ob = RESID ? a.O : a.O + bd.off;
    double mx = 0.0, my = 0.0;
    for (int r = 0; r < it.rows; r++) {
        const int y = y0 + r;
        R nn[4];
        ROW(y + 2 < ty ? y + 2 : ty - 1, nn);     // next iteration's U(y+1)
        const State<R> sn = make_state<M>(un, gam, g1);
        R Gn[4], Fw[4], Fe[4];
        hllc<M, false>(sc, sn, Gn);                // face y+1/2

        State<R> sw;
        sw.U0 = shfl_up1(sc.U0); sw.U1 = shfl_up1(sc.U1); sw.U2 = shfl_up1(sc.U2); sw.U3 = shfl_up1(sc.U3);
        sw.u = shfl_up1(sc.u); sw.v = shfl_up1(sc.v); sw.p = shfl_up1(sc.p); sw.E = shfl_up1(sc.E);
        sw.a = shfl_up1(sc.a);
        hllc<M, true>(sw, sc, Fw);                 // face x-1/2
        for (int k = 0; k < 4; k++) Fe[k] = shfl_dn1(Fw[k]);

        const R U[4] = {sc.U0, sc.U1, sc.U2, sc.U3};
        R Ln[4], Un[4];
        update_cell<M>(rdx, rdy, dtR, Fw, Fe, Gs, Gn, U, Ln, Un);
        if (out) {
            R* o = Ob + (int64_t)y * pitch + xraw;
            for (int k = 0; k < 4; k++) o[k * plane] = RESID ? Ln[k] : Un[k];
            if (!RESID) {
                double dx_, dy_;
                wave_speeds<M>(Un[0], Un[1], Un[2], Un[3], gam, g1, dx_, dy_);
                mx = dmax_nan(mx, dx_);
                my = dmax_nan(my, dy_);
            }
        }
        for (int k = 0; k < 4; k++) { Gs[k] = Gn[k]; un[k] = nn[k]; }
        sc = sn;
    }
    if (!RESID) store_maxima(imax_out, it.block, wave_max(mx), wave_max(my));
#undef ROW
}

// Persistent Godunov launch (no predictor, so never a fix-up).
template <class M, bool RESID>
__global__ __launch_bounds__(256) void k_godunov(const HfStepArgs<typename M::R> a)
{
    typedef typename M::R R;
    const int lane = threadIdx.x & 63;
    const unsigned nstatic = min(total_waves(), (unsigned)a.nitems);
    for (unsigned idx = first_ticket(a.sync, lane, nstatic);; idx = next_ticket(a.sync, lane, nstatic)) {
        if (idx >= (unsigned)a.nitems) break;
        const HfItem it = a.items[a.item0 + idx];
        double dt_step, dt_pred;
        if (a.use_fixed) {
            dt_step = a.fixed_step;
        } else {
            const HfBlockDev bd = a.blocks[it.block];
            long long kstep, kpred;
            step_dts<R>(a, it.block, bd, false, DT_SPEC, dt_step, dt_pred, kstep, kpred);
            const long long hs = kstep - a.hist_base;
            if (bd.problem == 0 && bd.lb == 0 && it.x0 == 0 && it.y0 == bd.g && hs >= 1 && hs <= a.hist_cap)
                a.hist[hs - 1] = dt_step;
        }
        godunov_item<M, RESID>(a, it, lane, dt_step, a.imax_new);
    }
}

// ------------------------------------------------------------- ghost fill --
// Mesh::apply_bcs (mesh/mesh/mesh.cpp:31-39) as a gather: the host replays the
// reference's sequential side order symbolically and records, for every
// ghost cell, which pre-fill cell (or Dirichlet constant) it ends up holding
// and which momentum components were negated on the way.  Reads come from S,
// writes go to D, so the fill is race-free in one launch.  Also reduces the
// ghosts' max wave speeds per block, and (step mode) advances each problem's
// TimeStep call counter and clears the partial maxima of the next step.

constexpr int GHOST_WAVES = 4;   // k_ghost workgroups: 256 threads
template <class M>
__global__ __launch_bounds__(256) void k_ghost(const HfGhost* __restrict__ tab, int n,
                                               const HfBlockDev* __restrict__ blocks,
                                               const typename M::R* __restrict__ S, typename M::R* __restrict__ D,
                                               const typename M::R* __restrict__ consts, unsigned long long* gmax,
                                               double gamma, GhostBook book)
{
    typedef typename M::R R;
    if (blockIdx.x == 0) {
        for (int p = threadIdx.x; p < book.nprob && book.bump; p += blockDim.x)
            book.calls[p] += (uint32_t)book.probs[p].cps;
        for (int i = threadIdx.x; i < book.nclear; i += blockDim.x) {
            if (book.clear0) book.clear0[i] = 0ull;
            if (book.clear1) book.clear1[i] = 0ull;
        }
        for (int i = threadIdx.x; i < book.nqueue; i += blockDim.x) book.queue[i] = 0u;
        for (int i = threadIdx.x; i < book.npending; i += blockDim.x) book.pending[i] = 0u;
        for (int b = threadIdx.x; b < book.nblocks && book.fixed; b += blockDim.x) {
            if (book.fixed[b]) {
                book.imax_prev[2 * b] = book.imax_fix[2 * b];
                book.imax_prev[2 * b + 1] = book.imax_fix[2 * b + 1];
            }
            book.skip[b] = book.skip_next[b];
            book.fixed[b] = 0u;
            book.imax_fix[2 * b] = 0ull;
            book.imax_fix[2 * b + 1] = 0ull;
        }
    }
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    double dx_ = 0.0, dy_ = 0.0;
    int blk = -1;
    if (gid < n) {
        const HfGhost e = tab[gid];
        R u[4];
        if (e.src_block >= 0) {
            const HfBlockDev sb = blocks[e.src_block];
            const R* p = S + sb.off + (e.src_cell & HF_CELL_MASK);
            load_cell<R>(p, sb.plane, u);
        } else {
            const R* c = consts + 4 * (-e.src_block - 1);
            u[0] = c[0]; u[1] = c[1]; u[2] = c[2]; u[3] = c[3];
        }
        if (e.src_cell & HF_FLIP_U) u[1] = -u[1];
        if (e.src_cell & HF_FLIP_V) u[2] = -u[2];
        const HfBlockDev db = blocks[e.dst_block];
        R* q = D + db.off + e.dst_cell;
        q[0] = u[0];
        q[db.plane] = u[1];
        q[2 * db.plane] = u[2];
        q[3 * db.plane] = u[3];
        double x, y;
        wave_speeds<M>(u[0], u[1], u[2], u[3], R(gamma), R(gamma - 1.0), x, y);
        dx_ = dmax_nan(0.0, x);
        dy_ = dmax_nan(0.0, y);
        blk = e.dst_block;
    }
    // The table is grouped by destination block: a wave covering one block
    // reduces to one value, and thread 0 merges consecutive waves of the same
    // block, so a workgroup issues one atomic per block it covers (device-
    // scope atomics on one address serialise: one per wave made this launch
    // take 20 us on 66 k ghost cells, one per workgroup 13 us).  A wave
    // spanning two blocks uses per-lane atomics.
    __shared__ int wblk[GHOST_WAVES];
    __shared__ double wmx[GHOST_WAVES], wmy[GHOST_WAVES];
    const int w = threadIdx.x >> 6;
    const int b0 = __shfl(blk, 0, 64);
    const bool uniform = __all(blk == b0 || blk < 0);
    if (uniform) {
        const double wx = wave_max(dx_), wy = wave_max(dy_);
        if (lane == 0) {
            wblk[w] = b0;
            wmx[w] = wx;
            wmy[w] = wy;
        }
    } else {
        if (blk >= 0) {
            atomicMax(&gmax[2 * blk], dbits(dx_));
            atomicMax(&gmax[2 * blk + 1], dbits(dy_));
        }
        if (lane == 0) wblk[w] = -1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int cb = -1;
        double cx = 0.0, cy = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
            if (wblk[k] == cb) {
                cx = (wmx[k] > cx) ? wmx[k] : cx;
                cy = (wmy[k] > cy) ? wmy[k] : cy;
                continue;
            }
            if (cb >= 0) {
                atomicMax(&gmax[2 * cb], dbits(cx));
                atomicMax(&gmax[2 * cb + 1], dbits(cy));
            }
            cb = wblk[k];
            cx = wmx[k];
            cy = wmy[k];
        }
        if (cb >= 0) {
            atomicMax(&gmax[2 * cb], dbits(cx));
            atomicMax(&gmax[2 * cb + 1], dbits(cy));
        }
    }
}

// -------------------------------------------------------------- wave scan --
// Max of |u|+c and |v|+c over cells of one block (TimeStep::operator(),
// timestep.cpp:37-55).  Interior cells are read from I, ghost cells from Gh;
// with interior_only the ghosts are skipped.
template <class M>
__global__ __launch_bounds__(256) void k_scan(HfBlockDev bd, const typename M::R* __restrict__ I,
                                              const typename M::R* __restrict__ Gh, int interior_only,
                                              double gamma, unsigned long long* outmax)
{
    typedef typename M::R R;
    const int64_t n = (int64_t)bd.tx * bd.ty;
    double mx = 0.0, my = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / bd.tx), x = (int)(i % bd.tx);
        const bool gh = x < bd.g || x >= bd.tx - bd.g || y < bd.g || y >= bd.ty - bd.g;
        if (gh && interior_only) continue;
        R u[4];
        load_cell<R>((gh ? Gh : I) + bd.off + (int64_t)y * bd.pitch + x, bd.plane, u);
        double dx_, dy_;
        wave_speeds<M>(u[0], u[1], u[2], u[3], R(gamma), R(gamma - 1.0), dx_, dy_);
        mx = dmax_nan(mx, dx_);
        my = dmax_nan(my, dy_);
    }
    mx = wave_max(mx);
    my = wave_max(my);
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&outmax[0], dbits(mx));
        atomicMax(&outmax[1], dbits(my));
    }
}

// --------------------------------------------------- conservation sums --
// Neumaier-compensated FP64 partial sums of the four conserved variables
// over a block's interior; one (sum, compensation) pair per workgroup and
// variable, combined on the host in workgroup order.
template <class R>
__global__ __launch_bounds__(256) void k_conserve(HfBlockDev bd, const R* __restrict__ A, double* part)
{
    __shared__ double ss[4][256];
    __shared__ double cc[4][256];
    double s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
    const int64_t n = (int64_t)bd.nx * bd.ny;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / bd.nx) + bd.g, x = (int)(i % bd.nx) + bd.g;
        const R* p = A + bd.off + (int64_t)y * bd.pitch + x;
        for (int k = 0; k < 4; k++) {
            const double v = (double)p[k * bd.plane];
            const double t = s[k] + v;
            c[k] += (__builtin_fabs(s[k]) >= __builtin_fabs(v)) ? ((s[k] - t) + v) : ((v - t) + s[k]);
            s[k] = t;
        }
    }
    for (int k = 0; k < 4; k++) { ss[k][threadIdx.x] = s[k]; cc[k][threadIdx.x] = c[k]; }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        double S = 0.0, Cm = 0.0;
        for (int i = 0; i < (int)blockDim.x; i++) {
            const double v = ss[k][i];
            const double t = S + v;
            Cm += (__builtin_fabs(S) >= __builtin_fabs(v)) ? ((S - t) + v) : ((v - t) + S);
            S = t;
            Cm += cc[k][i];
        }
        part[(blockIdx.x * 4 + k) * 2] = S;
        part[(blockIdx.x * 4 + k) * 2 + 1] = Cm;
    }
}

}  // namespace hf
