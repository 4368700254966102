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
// pass: wave shuffles, then a 64-bit atomic max per wave (store_maxima; where
// the host enables the probe, only from waves whose value raises the slot).
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

// Work items of a persistent launch: wave w first takes item w (no atomic, so
// a launch does not open with every wave hitting one counter), then tickets
// from the queue counter, offset by the number of waves.  Every
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
__device__ __forceinline__ unsigned next_ticket(uint32_t* sync, int lane, unsigned base)
{
    const unsigned t = atomicAdd(&sync[HF_SYNC_QUEUE], lane == 0 ? 1u : 0u);
    return __builtin_amdgcn_readfirstlane(t) + base;
}
// Tickets of a launch over items [0, n): wave w takes item w, then counter
// tickets.  When the items are no more than the waves, every wave takes at
// most its own item and no wave touches the counter: a wave would only learn
// there that the list is done, and those atomics all arrive together when the
// waves finish together and serialise at the counter.
__device__ __forceinline__ unsigned ticket_first(unsigned n)
{
    const unsigned w = wave_id();
    return (n <= total_waves() && w >= n) ? n : w;   // otherwise w < waves < n
}
__device__ __forceinline__ unsigned ticket_next(uint32_t* sync, int lane, unsigned n)
{
    const unsigned nw = total_waves();
    return n <= nw ? n : next_ticket(sync, lane, nw);
}

// What waves of one launch tell each other travels only through device-scope
// atomics (wave-speed maxima, fixed flags, the pending bitmap, the counters):
// writers update with atomic read-modify-writes (with the maxima probe, a
// writer skips one that a coherent load shows cannot raise the slot), and
// readers of values that decide results read with coherent_read
// (hf_physics.h), an atomic read-modify-write that changes nothing.  The new
// states themselves are not read inside the launch.

// One item of the launch is complete.  This wave's atomics (its maxima and
// fixed flag) must be performed before the count that waiting waves poll:
// s_waitcnt vmcnt(0) waits until the memory system has acknowledged every
// vector memory operation this wave issued (atomics without return
// included), then the count is raised.
__device__ __forceinline__ void item_done(uint32_t* sync, int lane, unsigned n = 1)
{
    __builtin_amdgcn_s_waitcnt(0xF70);   // vmcnt(0), expcnt and lgkmcnt unconstrained
    __hip_atomic_fetch_add(&sync[HF_SYNC_DONE], lane == 0 ? n : 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded wait until `target` items of the launch are complete (every item
// before position `target` of the queue: items are handed out in order and an
// item that waits never completes before the items it waits for).  The wait
// gives up after wait_ticks of the real-time counter (s_memrealtime, the
// clock HIP reports as hipDeviceAttributeWallClockRate), or at once when
// another wave has given up, and then returns false with the abort bit set:
// the caller leaves its item to the launch's last wave.  One load per poll
// (count and abort bit share a word); a load served late only delays the
// wait.
__device__ __forceinline__ bool wait_done(uint32_t* sync, unsigned target, unsigned long long wait_ticks)
{
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    for (;;) {
        const unsigned d = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(&sync[HF_SYNC_DONE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (d & HF_SYNC_ABORT_BIT) return false;
        if (d >= target) return true;
        if (__builtin_amdgcn_s_memrealtime() - t0 > wait_ticks) {
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

// Raises the slot at p to the wave-uniform bits b.  A slot is cleared before
// the launch that fills it and only grows during it, so a wave whose value
// does not exceed the slot's current one (a coherent load) leaves it alone:
// the same final maximum, and the device-scope atomics on one address, which
// serialise, come only from the waves that raise it.  The load puts a memory
// round trip at the end of the item: it pays where hundreds of items of one
// block end together (the 8-block channel: 4 % per step) and costs where few
// do (a 400^2 Godunov block: 15 %; 512 small blocks: 1.5 %), so the host
// enables it per context (HfStepArgs::probe_max).
__device__ __forceinline__ void raise_max(unsigned long long* p, unsigned long long b)
{
    const unsigned long long c = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)c);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(c >> 32));
    if ((((unsigned long long)hi << 32) | lo) < b) atomicMax(p, b);
}

// Per-wave maxima into the block's slots; every lane holds the reduced value.
// The value is passed as a scalar: the compiler then issues one atomic from
// one lane, instead of first reducing 64 lane values it cannot tell are equal
// (a 64-iteration scalar loop per atomic).
__device__ __forceinline__ void store_maxima(unsigned long long* out, int block, double mx, double my, bool probe)
{
    const unsigned long long bx = uniform_bits(mx), by = uniform_bits(my);
    if (probe) {
        raise_max(&out[2 * block], bx);
        raise_max(&out[2 * block + 1], by);
    } else {
        atomicMax(&out[2 * block], bx);
        atomicMax(&out[2 * block + 1], by);
    }
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
    if (!RESID) store_maxima(imax_out, it.block, wave_max(mx), wave_max(my), a.probe_max != 0);
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
    if (!RESID) store_maxima(imax_out, it.block, wave_max(mx), wave_max(my), a.probe_max != 0);
#undef COL
}

// The guess launch's plan of problem p (plan_scan with the main list's
// maxima): its first block to fix, the guess minimum m and the speculative
// minimum O.
template <class R>
__device__ __forceinline__ int guess_plan(const HfStepArgs<R>& a, int p, const HfProbDev& pr, double& O, double& m,
                                          long long& kstep)
{
    kstep = (long long)a.calls[p] - pr.cps + 1;
    double o = 1e16;
    const int end = pr.first + pr.nb;
    for (int c = pr.first + (int)(threadIdx.x & 63); c < end; c += 64) o = smin(o, term_old(a, c));
    O = wave_min(o);
    m = 0.0;
    return plan_scan<R, PRE_MAIN>(a, pr, a.nblocks, kstep, O, m, false);
}

// Check-launch rounds from round l on, while none of their blocks needs
// computing: round l holds the l-th block of every problem with more than l
// blocks, and its verdicts need the blocks before it final, which holds while
// every earlier round has been completed.  The queue counter jumps past each
// round passed over (its unissued items are counted complete).  Returns the
// first round to compute, or nrounds + 1.
template <class R>
__device__ int first_busy_round(const HfStepArgs<R>& a, int l, int lane, unsigned base)
{
    for (; l <= a.nrounds; l++) {
        bool need = false;
        for (int p = 0; p < a.nprob; p++) {
            const HfProbDev pr = a.probs[p];
            if (pr.nb <= l) continue;
            double O, mg, st, pd;
            long long ks, kp;
            guess_plan<R>(a, p, pr, O, mg, ks);
            need = check_verdict<R>(a, p, pr, a.nblocks, l, mg, st, pd, ks, kp) || need;
        }
        if (need) return l;
        const unsigned end = (unsigned)a.rounds[l + 1] - base;
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

// MUSCL-Hancock step with the reference-compatible dt (muscl_hancock.cpp:246
// with the sequential block loop of simulation.cpp:106-110: block l's
// predictor dt sees blocks 0..l-1 of its problem at U^{n+1}) in three
// persistent launches, each writing a block at most once (the XCDs' L2 caches
// are written back in no fixed order, so two writers of one cache line in one
// launch could leave either value):
//  1. main (k_muscl): every block except those marked skip[b] (the previous
//     step's verdict that the speculation fails), with the speculative
//     predictor dt (every block at U^n).  No block waits for another.
//  2. guess (k_muscl_guess): from the main list's maxima, each problem's
//     first wrong block L and, for every block l >= L, the exact dt of block
//     L scaled to l -- exact for l as long as blocks L..l-1 do not set a new
//     minimum.  Every block whose dt differs from its speculative one (or
//     that was skipped) is computed with it.  No waits.
//  3. check (k_muscl_check): from the final maxima, the first block whose dt
//     so far is not the exact one; normally none, and the launch returns at
//     once.  Otherwise rounds follow from that block on: round l holds block
//     l of every problem and waits (bounded) until every earlier round is
//     complete, then recomputes the blocks whose dt is still wrong.
// Items [0, nmain) form the main list, [nmain, nitems) the rounds: round l
// lists the l-th block of every problem (the guess and check launches take
// them in that order).  Without a CFL dt (use_fixed) or in uniform mode the
// main launch is the whole step.
// Exclusive prefix sum over the wave's lanes; total: the sum over all lanes.
__device__ __forceinline__ unsigned lane_prefix(unsigned x, unsigned& total)
{
    const int lane = threadIdx.x & 63;
    unsigned v = x;
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    total = __builtin_amdgcn_readlane(v, 63);
    return v - x;
}

// A ticket of a launch that takes only some of its item ranges: the ranges
// are visited 64 at a time (lane <-> range), each lane holding its range's
// first item, its count (0: not taken) and the items of the taken ranges
// before it in the chunk.  Tickets count taken items only, so no ticket is
// spent on an item that is not computed; a wave's tickets only grow, so its
// cursor only moves forward.
struct RangeCursor {
    int c0;            // first range of the current chunk
    unsigned base;     // taken items before the chunk
    unsigned total;    // taken items in the chunk
    int first;         // this lane's range: first item
    unsigned cnt, pre; // this lane's range: taken items, taken items before it in the chunk
};

// Item position of ticket t in the chunk that holds it: the lane whose range
// covers it, and the item.
__device__ __forceinline__ unsigned cursor_item(const RangeCursor& cur, unsigned t, int& lane_of)
{
    const unsigned r = t - cur.base;
    const unsigned long long bal = __ballot(cur.cnt > 0u && r >= cur.pre && r < cur.pre + cur.cnt);
    if (!bal) return 0xffffffffu;   // not reached: the chunk's ranges cover [base, base + total)
    lane_of = __builtin_ctzll(bal);
    const unsigned pre = __builtin_amdgcn_readlane(cur.pre, lane_of);
    const int first = __builtin_amdgcn_readlane(cur.first, lane_of);
    return (unsigned)first + (r - pre);
}

// Main list with skipped blocks: ranges are the (phase, block) segments in
// list order, a segment taken unless its block is skipped.
template <class R>
__device__ __forceinline__ void main_chunk(const HfStepArgs<R>& a, RangeCursor& cur, int c0)
{
    const int k = c0 + (int)(threadIdx.x & 63);
    const int nseg = 3 * a.nblocks;
    cur.c0 = c0;
    cur.first = 0;
    cur.cnt = 0u;
    if (k < nseg && !a.skip[k % a.nblocks]) {
        cur.first = a.segs[2 * k];
        cur.cnt = (unsigned)a.segs[2 * k + 1];
    }
    cur.pre = lane_prefix(cur.cnt, cur.total);
}

template <class M, int LIM, bool RESID>
__device__ __forceinline__ void main_item(const HfStepArgs<typename M::R>& a, unsigned idx, int lane, int& cached,
                                          double& c_step, double& c_pred, long long& c_kstep, long long& c_kpred)
{
    typedef typename M::R R;
    const HfItem it = a.items[a.item0 + idx];
    if (a.use_fixed) {
        muscl_go<M, LIM, RESID>(a, it, lane, a.fixed_step, a.fixed_pred, a.imax_new);
        return;
    }
    if (a.skip[it.block]) return;   // left to the guess launch
    const HfBlockDev bd = a.blocks[it.block];
    if (it.block != cached) {
        cached = it.block;
        step_dts<R>(a, bd, true, c_step, c_pred, c_kstep, c_kpred);
    }
    record_dts<R>(a, it, bd, true, c_step, c_pred, c_kstep, c_kpred);
    muscl_go<M, LIM, RESID>(a, it, lane, c_step, c_pred, a.imax_new);
}

template <class M, int LIM, bool RESID>
__global__ __launch_bounds__(256) void k_muscl(const HfStepArgs<typename M::R> a)
{
    const int lane = threadIdx.x & 63;
    int cached = -1;   // the block whose speculative dts this wave has computed
    double c_step = 0.0, c_pred = 0.0;
    long long c_kstep = 0, c_kpred = 0;
    const unsigned nsk = (a.nskip && !a.use_fixed) ? __builtin_amdgcn_readfirstlane(*a.nskip) : 0u;
    if (nsk == 0u) {   // every item of the list
        const unsigned n = (unsigned)a.nmain;
        for (unsigned idx = ticket_first(n); idx < n; idx = ticket_next(a.sync, lane, n))
            main_item<M, LIM, RESID>(a, idx, lane, cached, c_step, c_pred, c_kstep, c_kpred);
        return;
    }
    // tickets over the segments of the blocks not skipped
    const int nseg = 3 * a.nblocks;
    unsigned n = 0u;
    for (int c0 = 0; c0 < nseg; c0 += 64) {
        const int k = c0 + lane;
        const unsigned cnt = (k < nseg && !a.skip[k % a.nblocks]) ? (unsigned)a.segs[2 * k + 1] : 0u;
        unsigned tot;
        lane_prefix(cnt, tot);
        n += tot;
    }
    RangeCursor cur;
    main_chunk<typename M::R>(a, cur, 0);
    cur.base = 0u;
    for (unsigned t = ticket_first(n); t < n; t = ticket_next(a.sync, lane, n)) {
        while (t >= cur.base + cur.total) {
            const unsigned nb_ = cur.base + cur.total;
            if (cur.c0 + 64 >= nseg) return;   // not reached: the chunks hold n items
            main_chunk<typename M::R>(a, cur, cur.c0 + 64);
            cur.base = nb_;
        }
        int j;
        const unsigned idx = cursor_item(cur, t, j);
        if (idx >= (unsigned)a.nmain) return;
        main_item<M, LIM, RESID>(a, idx, lane, cached, c_step, c_pred, c_kstep, c_kpred);
    }
}

// The guess launch's view of problem p: its plan and the cursor over its
// blocks from the plan's first block on (a range per block: its items in the
// fix-up rounds, taken if the block needs computing).
struct GuessProblem {
    int p, first, nb;
    double O, m;
    long long kstep;
};

template <class R>
__device__ __forceinline__ bool guess_need(const HfStepArgs<R>& a, const GuessProblem& g, int c, int b)
{
    const double cf = cfl_eff(a.cfl, g.kstep + 1 + c);
    return a.skip[b] != 0u || cf * g.O != cf * smin(1e16, g.m);
}

template <class R>
__device__ __forceinline__ void guess_chunk(const HfStepArgs<R>& a, const GuessProblem& g, RangeCursor& cur, int c0)
{
    const int c = c0 + (int)(threadIdx.x & 63);
    cur.c0 = c0;
    cur.first = 0;
    cur.cnt = 0u;
    if (c >= g.first && c < g.nb) {
        const int b = a.probs[g.p].first + c;
        if (guess_need<R>(a, g, c, b)) {
            cur.first = a.fixrange[2 * b];
            cur.cnt = (unsigned)a.fixrange[2 * b + 1];
        }
    }
    cur.pre = lane_prefix(cur.cnt, cur.total);
}

// The next problem from p on with blocks to compute (nprob if none).
template <class R>
__device__ __forceinline__ void guess_enter(const HfStepArgs<R>& a, GuessProblem& g, int p)
{
    for (; p < a.nprob; p++) {
        const HfProbDev pr = a.probs[p];
        if (pr.nb < 2) continue;
        g.p = p;
        g.nb = pr.nb;
        g.first = guess_plan<R>(a, p, pr, g.O, g.m, g.kstep);
        if (g.first < pr.nb) return;
    }
    g.p = a.nprob;
}

template <class M, int LIM>
__global__ __launch_bounds__(256) void k_muscl_guess(const HfStepArgs<typename M::R> a)
{
    typedef typename M::R R;
    const int lane = threadIdx.x & 63;
    // items to compute, over every problem; the walk below starts from the
    // first problem's first chunk, kept from this pass
    unsigned n = 0u;
    GuessProblem g, g0;
    RangeCursor cur;
    g0.p = a.nprob;
    for (guess_enter<R>(a, g, 0); g.p < a.nprob; guess_enter<R>(a, g, g.p + 1))
        for (int c0 = g.first & ~63; c0 < g.nb; c0 += 64) {
            RangeCursor ck;
            guess_chunk<R>(a, g, ck, c0);
            if (g0.p == a.nprob) {
                g0 = g;
                cur = ck;
            }
            n += ck.total;
        }
    if (n == 0u) return;
    g = g0;
    cur.base = 0u;
    int cb = -1;   // the block whose dts this wave holds
    double dt_step = 0.0, dt_pred = 0.0;
    long long kpred = 0;
    for (unsigned t = ticket_first(n); t < n; t = ticket_next(a.sync, lane, n)) {
        while (t >= cur.base + cur.total) {
            const unsigned nb_ = cur.base + cur.total;
            int c0 = cur.c0 + 64;
            if (c0 >= g.nb) {
                guess_enter<R>(a, g, g.p + 1);
                if (g.p >= a.nprob) return;   // not reached: the problems hold n items
                c0 = g.first & ~63;
            }
            guess_chunk<R>(a, g, cur, c0);
            cur.base = nb_;
        }
        int j;
        const unsigned idx = cursor_item(cur, t, j);
        if (idx < (unsigned)a.nmain || idx >= (unsigned)a.nitems) return;
        const HfItem it = a.items[a.item0 + idx];
        const HfBlockDev bd = a.blocks[it.block];
        if (it.block != cb) {
            cb = it.block;
            kpred = g.kstep + 1 + bd.lb;
            dt_step = cfl_eff(a.cfl, g.kstep) * g.O;
            dt_pred = cfl_eff(a.cfl, kpred) * smin(1e16, g.m);
            a.fixed[cb] = 1u;   // the same value from every wave
        }
        record_dts<R>(a, it, bd, false, dt_step, dt_pred, g.kstep, kpred);
        muscl_go<M, LIM, false>(a, it, lane, dt_step, dt_pred, a.imax_fix);
    }
}

template <class M, int LIM>
__global__ __launch_bounds__(256) void k_muscl_check(const HfStepArgs<typename M::R> a)
{
    typedef typename M::R R;
    const int lane = threadIdx.x & 63;
    // the first round with a wrong block; wave 0 records the skip verdicts
    int first = a.nrounds + 1;
    const bool rec = wave_id() == 0;
    for (int p = 0; p < a.nprob; p++) {
        const HfProbDev pr = a.probs[p];
        if (pr.nb < 2) continue;
        double O, m;
        long long kstep;
        guess_plan<R>(a, p, pr, O, m, kstep);
        const int w = plan_scan<R, PRE_GUESS>(a, pr, a.nblocks, kstep, O, m, rec);
        first = w < first ? w : first;
    }
    if (first > a.nrounds) return;   // every block has its exact dt

    unsigned long long* imax_chk = a.imax_fix + 2 * a.nblocks;
    uint32_t* fixed_chk = a.fixed + a.nblocks;
    const unsigned base = (unsigned)a.rounds[first];   // item of queue counter value 0
    int cached = -1, cp = -1;   // the block whose verdict this wave holds, the problem of its guess minimum
    bool c_need = false;
    double cm = 0.0, c_step = 0.0, c_pred = 0.0;
    long long c_kstep = 0, c_kpred = 0;
    int r_from = -1, r_busy = 0;   // first_busy_round(r_from) as this wave last computed it
    for (unsigned idx = next_ticket(a.sync, lane, base);; idx = next_ticket(a.sync, lane, base)) {
        if (idx >= (unsigned)a.nitems) break;
        const HfItem it = a.items[a.item0 + idx];
        const int l = round_of<R>(a, idx);
        if (l != r_from) {
            if (!wait_done(a.sync, (unsigned)a.rounds[l] - base, a.wait_ticks)) {
                const unsigned k = idx - (unsigned)a.nmain;
                atomicOr(&a.pending[k >> 5], 1u << (k & 31));
                continue;   // left to the last wave, not counted complete
            }
            r_from = l;
            r_busy = first_busy_round<R>(a, l, lane, base);
        }
        if (r_busy == l) {
            const HfBlockDev bd = a.blocks[it.block];
            if (it.block != cached) {
                cached = it.block;
                const HfProbDev pr = a.probs[bd.problem];
                if (bd.problem != cp) {
                    cp = bd.problem;
                    double O;
                    long long ks;
                    guess_plan<R>(a, cp, pr, O, cm, ks);
                }
                c_need = check_verdict<R>(a, bd.problem, pr, a.nblocks, bd.lb, cm, c_step, c_pred, c_kstep, c_kpred);
            }
            if (c_need) {
                atomicExch(&fixed_chk[it.block], 1u);
                record_dts<R>(a, it, bd, false, c_step, c_pred, c_kstep, c_kpred);
                muscl_go<M, LIM, false>(a, it, lane, c_step, c_pred, imax_chk);
            }
        }
        item_done(a.sync, lane);
    }
    // The last wave to leave computes the items whose wait gave up, in queue
    // order; every other wave has left, so all it reads is final, and none of
    // those items was written by another wave.
    __builtin_amdgcn_s_waitcnt(0xF70);   // this wave's atomics are performed before it counts as gone
    const unsigned left = __builtin_amdgcn_readfirstlane(atomicAdd(&a.sync[HF_SYNC_EXITED], lane == 0 ? 1u : 0u));
    if (left != total_waves() - 1) return;
    if (!(__builtin_amdgcn_readfirstlane(coherent_read(&a.sync[HF_SYNC_DONE])) & HF_SYNC_ABORT_BIT)) return;
    const unsigned nfix = (unsigned)(a.nitems - a.nmain);
    for (unsigned w = 0; w * 32 < nfix; w++) {
        const unsigned bits = __builtin_amdgcn_readfirstlane(coherent_read(&a.pending[w]));
        for (unsigned k = 0; k < 32; k++) {
            if (!((bits >> k) & 1u)) continue;
            const unsigned idx = a.nmain + w * 32 + k;
            const HfItem it = a.items[a.item0 + idx];
            const HfBlockDev bd = a.blocks[it.block];
            const HfProbDev pr = a.probs[bd.problem];
            double O, mg, dt_step, dt_pred;
            long long kstep, kpred;
            guess_plan<R>(a, bd.problem, pr, O, mg, kstep);
            if (check_verdict<R>(a, bd.problem, pr, a.nblocks, bd.lb, mg, dt_step, dt_pred, kstep, kpred)) {
                atomicExch(&fixed_chk[it.block], 1u);
                record_dts<R>(a, it, bd, false, dt_step, dt_pred, kstep, kpred);
                muscl_go<M, LIM, false>(a, it, lane, dt_step, dt_pred, imax_chk);
                __builtin_amdgcn_s_waitcnt(0xF70);   // its maxima are performed before the next verdict reads them
            }
        }
        atomicExch(&a.pending[w], 0u);
    }
}

// ---------------------------------------------------------------- Godunov --
// GodunovScheme::operator() (scheme/godunov/godunov.cpp:27-118) fused with
// the forward Euler update.  Each cell's primitives and sound speed are
// computed once and shared by its four faces.  Output columns per wave: 62.
// dt_of() gives the step's dt; it is called after the item's first rows are
// requested and the first face flux computed, none of which needs the dt,
// so the dt's chain of dependent loads overlaps their memory latency.
template <class M, bool RESID, class DtOf>
__device__ __forceinline__ void godunov_item(const HfStepArgs<typename M::R>& a, const HfItem it, const int lane,
                                             DtOf&& dt_of, unsigned long long* __restrict__ imax_out)
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
    const R gam = R(a.gamma), g1 = R(a.gamma - 1.0);

    const int y0 = it.y0;
    R ub[4], uc[4], un[4];
    ROW(y0 - 1, ub);
    ROW(y0, uc);
    ROW(y0 + 1, un);                               // loaded one iteration ahead
    State<R> sb = make_state<M>(ub, gam, g1);
    State<R> sc = make_state<M>(uc, gam, g1);
    R Gs[4];
    hllc<M, false>(sb, sc, Gs);                    // face y0-1/2
    const R dtR = R(dt_of());
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
    if (!RESID) store_maxima(imax_out, it.block, wave_max(mx), wave_max(my), false);
#undef ROW
}

// Persistent Godunov launch (no predictor, so never a fix-up).
template <class M, bool RESID>
__global__ __launch_bounds__(256) void k_godunov(const HfStepArgs<typename M::R> a)
{
    typedef typename M::R R;
    const int lane = threadIdx.x & 63;
    const unsigned n = (unsigned)a.nitems;
    for (unsigned idx = ticket_first(n); idx < n; idx = ticket_next(a.sync, lane, n)) {
        const HfItem it = a.items[a.item0 + idx];
        auto dt_of = [&]() -> double {
            if (a.use_fixed) return a.fixed_step;
            const HfBlockDev bd = a.blocks[it.block];
            double dt_step, dt_pred;
            long long kstep, kpred;
            step_dts<R>(a, bd, false, dt_step, dt_pred, kstep, kpred);
            const long long hs = kstep - a.hist_base;
            if (bd.problem == 0 && bd.lb == 0 && it.x0 == 0 && it.y0 == bd.g && hs >= 1 && hs <= a.hist_cap)
                a.hist[hs - 1] = dt_step;
            return dt_step;
        };
        godunov_item<M, RESID>(a, it, lane, dt_of, a.imax_new);
    }
}

// ------------------------------------------------------------- ghost fill --
// Mesh::apply_bcs (mesh/mesh/mesh.cpp:31-39) as a gather: the host replays the
// reference's sequential side order symbolically and records, for every
// ghost cell, which pre-fill cell (or Dirichlet constant) it ends up holding
// and which momentum components were negated on the way.  Reads come from S,
// writes go to D, so the fill is race-free in one launch.  Also reduces the
// ghosts' max wave speeds per block, and (step mode, in one extra workgroup
// after those of the ghost cells) advances each problem's TimeStep call
// counter and clears the partial maxima of the next step.

constexpr int GHOST_WAVES = 4;   // k_ghost workgroups: 256 threads
template <class M>
__global__ __launch_bounds__(256) void k_ghost(const HfGhost* __restrict__ tab, int n,
                                               const HfBlockDev* __restrict__ blocks,
                                               const typename M::R* __restrict__ S, typename M::R* __restrict__ D,
                                               const typename M::R* __restrict__ consts, unsigned long long* gmax,
                                               double gamma, GhostBook book)
{
    typedef typename M::R R;
    // The last workgroup keeps the step's books (none of it reads or writes a
    // ghost cell or gmax), beside the ghost cells instead of ahead of some.
    if (blockIdx.x == gridDim.x - 1) {
        for (int p = threadIdx.x; p < book.nprob && book.bump; p += blockDim.x)
            book.calls[p] += (uint32_t)book.probs[p].cps;
        for (int i = threadIdx.x; i < book.nclear; i += blockDim.x) {
            if (book.clear0) book.clear0[i] = 0ull;
            if (book.clear1) book.clear1[i] = 0ull;
        }
        for (int i = threadIdx.x; i < book.nqueue; i += blockDim.x) book.queue[i] = 0u;
        for (int i = threadIdx.x; i < book.npending; i += blockDim.x) book.pending[i] = 0u;
        __syncthreads();   // the skip count below starts from zero
        const int nb = book.nblocks;
        if (book.fixstats && book.fixed && threadIdx.x == 0) atomicAdd(&book.fixstats[0], 1ull);
        for (int b = threadIdx.x; b < nb && book.fixed; b += blockDim.x) {
            if (book.fixstats) {
                if (book.skip[b]) atomicAdd(&book.fixstats[1], 1ull);
                if (book.fixed[b]) atomicAdd(&book.fixstats[2], 1ull);
                if (book.fixed[nb + b]) atomicAdd(&book.fixstats[3], 1ull);
            }
            // the check launch's recomputation over the guess launch's
            const int s = book.fixed[nb + b] ? 1 : book.fixed[b] ? 0 : -1;
            if (s >= 0) {
                book.imax_prev[2 * b] = book.imax_fix[2 * (s * nb + b)];
                book.imax_prev[2 * b + 1] = book.imax_fix[2 * (s * nb + b) + 1];
            }
            book.skip[b] = book.skip_next[b];
            if (book.skip_next[b]) atomicAdd(&book.queue[HF_SLOT_MAIN * HF_SYNC_WORDS + HF_SYNC_NSKIP], 1u);
            for (int t = 0; t < 2; t++) {
                book.fixed[t * nb + b] = 0u;
                book.imax_fix[2 * (t * nb + b)] = 0ull;
                book.imax_fix[2 * (t * nb + b) + 1] = 0ull;
            }
        }
        return;
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
    // take 20 us on 66 k ghost cells, one per workgroup 13 us; reading the
    // slot first, as raise_max does, made it slower still).  A wave spanning
    // two blocks uses per-lane atomics.
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
