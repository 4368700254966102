// hf_capi.cpp -- the C ABI (include/hyperflow.h) over the gfx950 kernels.
//
// A context owns two state buffers (ping-pong).  buf[cur] always holds the
// complete current state (interior and ghosts).  A step fills the ghosts of
// buf[cur^1] from buf[cur] (table-driven gather, the reference's side order
// replayed on the host), then the update kernels read the interior from
// buf[cur] and the ghosts from buf[cur^1] and write the new interior into
// buf[cur^1]; cur flips.  The dt of every step stays on the device: the
// update kernels derive it from per-block wave-speed maxima that the previous
// update and the ghost fill reduced, and from a device-resident TimeStep call
// counter that the ghost fill advances.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/hyperflow.h"
#include "../../include/hyperflow_learned.h"
#include "hf_common.h"
#include "hf_context.h"
#include "hf_launch.h"
#include "hf_noalloc.h"

thread_local std::string g_err;

int hf_fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

namespace {
std::atomic<bool> g_exiting{false};
// Registered when the library is loaded, i.e. after the HIP runtime it links
// against: atexit handlers run in reverse order, so this one runs first.
const int g_exit_hook = std::atexit([] { g_exiting.store(true); });
}  // namespace

bool hf_process_exiting() { return g_exiting.load(); }

namespace {

inline int fail(int code, const std::string& msg) { return hf_fail(code, msg); }

inline double smin(double a, double b) { return (b < a) ? b : a; }
inline double bitsd(unsigned long long b)
{
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

// Symbolic value of a cell during the host replay of Mesh::apply_bcs: the
// pre-fill cell (or Dirichlet constant) it holds, and momentum negations.
struct Prov {
    int32_t blk;    // source block, or -1 for a constant
    int32_t x, y;   // source cell
    int32_t cidx;   // constant index (block*4 + side)
    uint32_t flip;  // HF_FLIP_U | HF_FLIP_V
};

}  // namespace


namespace {

// ------------------------------------------------------------ validation --
int validate(const hf_case_desc* d)
{
    if (!d || !d->blocks || d->nblocks < 1) return fail(HF_EINVAL, "case: no blocks");
    if (d->scheme != HF_GODUNOV && d->scheme != HF_MUSCL_HANCOCK) return fail(HF_EINVAL, "case: unknown scheme");
    if (d->scheme == HF_MUSCL_HANCOCK && d->limiter != HF_MINBEE && d->limiter != HF_VANLEER)
        return fail(HF_EINVAL, "case: unknown limiter");
    if (d->precision != HF_FP64 && d->precision != HF_FP32) return fail(HF_EINVAL, "case: unknown precision");
    if (d->dt_mode != HF_DT_COMPAT && d->dt_mode != HF_DT_UNIFORM) return fail(HF_EINVAL, "case: unknown dt mode");
    if (!std::isfinite(d->gamma) || !std::isfinite(d->cfl) || d->cfl <= 0.0)
        return fail(HF_EINVAL, "case: gamma and cfl must be finite, cfl > 0");
    if (d->tile_rows < 0 || d->tail_rows < 0 || d->fix_rows < 0)
        return fail(HF_EINVAL, "case: tile_rows, tail_rows and fix_rows must be >= 0");
    const int g = d->scheme == HF_MUSCL_HANCOCK ? 2 : 1;
    int prev_problem = 0;
    for (int b = 0; b < d->nblocks; b++) {
        const hf_block_desc& B = d->blocks[b];
        char where[64];
        std::snprintf(where, sizeof where, "block %d: ", b);
        if (B.nx < 1 || B.ny < 1) return fail(HF_EINVAL, std::string(where) + "needs nx, ny >= 1");
        if (!(B.x_right > B.x_left) || !(B.y_top > B.y_bottom) || !std::isfinite(B.x_right - B.x_left) ||
            !std::isfinite(B.y_top - B.y_bottom))
            return fail(HF_EINVAL, std::string(where) + "extents must be finite with right > left, top > bottom");
        const int64_t tx = B.nx + 2 * g, ty = B.ny + 2 * g;
        if ((tx + 64) * ty >= (int64_t)HF_CELL_MASK) return fail(HF_EINVAL, std::string(where) + "too many cells");
        if (B.problem < prev_problem || B.problem > prev_problem + 1 || (b == 0 && B.problem != 0))
            return fail(HF_EINVAL, std::string(where) + "problems must be numbered 0,1,.. over contiguous blocks");
        prev_problem = B.problem;
        for (int s = 0; s < 4; s++) {
            const int k = B.bc[s];
            if (k < HF_TRANSMISSIVE || k > HF_INTERNAL) return fail(HF_EINVAL, std::string(where) + "unknown bc");
            if (k == HF_INTERNAL) {
                const int n = B.neighbour[s];
                if (n < 0 || n >= d->nblocks) return fail(HF_EINVAL, std::string(where) + "Internal neighbour out of range");
                const hf_block_desc& E = d->blocks[n];
                if (E.problem != B.problem)
                    return fail(HF_EINVAL, std::string(where) + "Internal neighbour in another problem");
                if ((s == HF_LEFT || s == HF_RIGHT) ? (E.ny != B.ny) : (E.nx != B.nx))
                    return fail(HF_EINVAL, std::string(where) + "Internal neighbour must match the shared edge's cell count");
            }
            if (k == HF_DIRICHLET)
                for (int v = 0; v < 4; v++)
                    if (!std::isfinite(B.state[s][v])) return fail(HF_EINVAL, std::string(where) + "non-finite Dirichlet state");
        }
    }
    return HF_OK;
}

// ------------------------------------------------ ghost table (host replay) --
// Replays Mesh::apply_bcs (mesh.cpp:31-39) on provenance records instead of
// values.  The side loops follow Transmissive (transmissive_boundary.cpp:25-104),
// Reflective (reflective_boundary.cpp:25-120), Dirichlet
// (dirichlet_boundary.cpp:25-104) and Internal (internal_boundary.cpp:27-108):
// every side writes its g layers over the block's full padded extent.
struct Replay {
    const hf_ctx& c;
    std::vector<std::vector<Prov>> prov;   // per block, ghost cells in perimeter order

    explicit Replay(const hf_ctx& ctx) : c(ctx), prov(ctx.nb)
    {
        for (int b = 0; b < c.nb; b++) {
            const HfBlockDev& B = c.blocks[b];
            prov[b].resize(count(B));
            for (int y = 0; y < B.ty; y++)
                for (int x = 0; x < B.tx; x++)
                    if (ghost(B, x, y)) prov[b][index(B, x, y)] = Prov{b, x, y, -1, 0u};
        }
    }
    static bool ghost(const HfBlockDev& B, int x, int y)
    {
        return x < B.g || x >= B.tx - B.g || y < B.g || y >= B.ty - B.g;
    }
    static int64_t count(const HfBlockDev& B) { return 2LL * B.g * B.tx + 2LL * B.g * (B.ty - 2 * B.g); }
    // perimeter index: bottom band, top band, then left/right of the middle rows
    static int64_t index(const HfBlockDev& B, int x, int y)
    {
        const int g = B.g;
        if (y < g) return (int64_t)y * B.tx + x;
        if (y >= B.ty - g) return (int64_t)g * B.tx + (int64_t)(y - (B.ty - g)) * B.tx + x;
        const int64_t base = 2LL * g * B.tx + (int64_t)(y - g) * 2 * g;
        return x < g ? base + x : base + g + (x - (B.tx - g));
    }
    Prov read(int b, int x, int y) const
    {
        const HfBlockDev& B = c.blocks[b];
        if (ghost(B, x, y)) return prov[b][index(B, x, y)];
        return Prov{b, x, y, -1, 0u};
    }
    void write(int b, int x, int y, const Prov& p) { prov[b][index(c.blocks[b], x, y)] = p; }

    void side(int b, int s)
    {
        const HfBlockDev& B = c.blocks[b];
        const hf_block_desc& D = c.bdesc[b];
        const int g = B.g, tx = B.tx, ty = B.ty, kind = D.bc[s];
        const int e = kind == HF_INTERNAL ? D.neighbour[s] : -1;
        const HfBlockDev* E = e >= 0 ? &c.blocks[e] : nullptr;
        const int n = (s == HF_BOTTOM || s == HF_TOP) ? tx : ty;
        const uint32_t negate = (s == HF_BOTTOM || s == HF_TOP) ? HF_FLIP_V : HF_FLIP_U;
        for (int k = 0; k < n; k++) {
            int gx[2], gy[2], sx[2], sy[2];
            for (int l = 0; l < g; l++) {   // l = 0: layer next to the interior
                switch (s) {
                case HF_BOTTOM:
                    gx[l] = k; gy[l] = g - 1 - l;
                    sx[l] = k; sy[l] = E ? E->ty - g - 1 - l : g + l;
                    break;
                case HF_RIGHT:
                    gx[l] = tx - g + l; gy[l] = k;
                    sx[l] = E ? g + l : tx - g - 1 - l; sy[l] = k;
                    break;
                case HF_TOP:
                    gx[l] = k; gy[l] = ty - g + l;
                    sx[l] = k; sy[l] = E ? g + l : ty - g - 1 - l;
                    break;
                default:
                    gx[l] = g - 1 - l; gy[l] = k;
                    sx[l] = E ? E->tx - g - 1 - l : g + l; sy[l] = k;
                    break;
                }
            }
            // Internal right/top write the outer layer first (internal_boundary.cpp:62-63,104-105)
            int order[2] = {0, 1};
            if (kind == HF_INTERNAL && g == 2 && (s == HF_RIGHT || s == HF_TOP)) { order[0] = 1; order[1] = 0; }
            for (int q = 0; q < g; q++) {
                const int l = order[q];
                Prov p;
                if (kind == HF_DIRICHLET) p = Prov{-1, 0, 0, b * 4 + s, 0u};
                else p = read(E ? e : b, sx[l], sy[l]);
                write(b, gx[l], gy[l], p);
            }
            if (kind == HF_REFLECTIVE)
                for (int q = 0; q < g; q++) {
                    Prov p = read(b, gx[q], gy[q]);
                    p.flip ^= negate;
                    write(b, gx[q], gy[q], p);
                }
        }
    }
};

void build_tables(const hf_ctx& c, std::vector<HfGhost>& fill, std::vector<HfGhost>& copy)
{
    Replay r(c);
    for (int b = 0; b < c.nb; b++)
        for (int s = 0; s < 4; s++) r.side(b, s);
    for (int b = 0; b < c.nb; b++) {
        const HfBlockDev& B = c.blocks[b];
        // emit in memory order of the destination for coalescing
        for (int y = 0; y < B.ty; y++)
            for (int x = 0; x < B.tx; x++) {
                if (!Replay::ghost(B, x, y)) continue;
                const Prov& p = r.prov[b][Replay::index(B, x, y)];
                HfGhost e;
                e.dst_block = b;
                e.dst_cell = (uint32_t)((int64_t)y * B.pitch + x);
                if (p.blk >= 0) {
                    e.src_block = p.blk;
                    e.src_cell = (uint32_t)((int64_t)p.y * c.blocks[p.blk].pitch + p.x) | p.flip;
                } else {
                    e.src_block = -(p.cidx + 1);
                    e.src_cell = p.flip;
                }
                fill.push_back(e);
                HfGhost id;
                id.dst_block = b;
                id.dst_cell = e.dst_cell;
                id.src_block = b;
                id.src_cell = e.dst_cell;
                copy.push_back(id);
            }
    }
}

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// ---------------------------------------------------------------- launches --
// Offsets of the main-list segments and the fix-up ranges in plan_table().
size_t segs_offset(const hf_ctx& c) { return (size_t)c.nrounds + 2; }
size_t fixrange_offset(const hf_ctx& c) { return segs_offset(c) + c.segs.size(); }

template <class R>
HfStepArgs<R> step_args(hf_ctx* c, int item0, int nitems, const void* A, const void* G, void* O)
{
    HfStepArgs<R> a;
    std::memset(&a, 0, sizeof a);
    a.blocks = c->d_blocks;
    a.probs = c->d_probs;
    a.items = c->d_items;
    a.item0 = item0;
    a.nitems = nitems;
    a.A = (const R*)A;
    a.G = (const R*)G;
    a.O = (R*)O;
    a.nprob = c->np;
    a.nblocks = c->nb;
    a.nmain = nitems;
    a.sync = c->d_sync;
    a.rounds = c->d_rounds;
    a.segs = c->d_rounds + segs_offset(*c);
    a.fixrange = c->d_rounds + fixrange_offset(*c);
    a.nskip = nullptr;
    a.wait_ticks = c->wait_ticks;
    a.probe_max = c->probe_max;
    a.nrounds = 0;
    a.pending = c->d_pending;
    a.npending = c->npending;
    a.imax_fix = c->d_imax_fix;
    a.fixed = c->d_fixed;
    a.skip = c->d_skip;
    a.skip_next = c->d_skip_next;
    a.imax_old = c->imax(c->cur);
    a.imax_new = c->imax(c->cur ^ 1);
    a.gmax = c->gmax(c->cur);
    a.calls = c->d_calls;
    a.hist = c->d_hist;
    a.hist_base = c->hist_base;
    a.hist_cap = c->hist_cap;
    a.dt_mode = c->dt_mode;
    a.gamma = c->gamma;
    a.cfl = c->cfl;
    return a;
}

// A pair of timing events from the context's pool; the pending list takes
// them once the launches that stamp them are queued.
int take_events(hf_ctx* c, hipEvent_t& e0, hipEvent_t& e1)
{
    if (c->ev_free.empty()) {
        HF_HIP(hipEventCreate(&e0));
        HF_HIP(hipEventCreate(&e1));
    } else {
        e0 = c->ev_free.back().first;
        e1 = c->ev_free.back().second;
        c->ev_free.pop_back();
    }
    return HF_OK;
}

// One persistent update launch of kind `kind` (main list, guess or check)
// over items [item0, item0+nitems), taking tickets with the sync counters of
// slot `slot` (zeroed before the launch runs); the items after the first
// nmain are the fix-up rounds.  e0 / e1 (optional) are stamped at the
// dispatch's start / end.
template <class R>
int launch_update_r(hf_ctx* c, int kind, int item0, int nitems, int nmain, const void* A, const void* G, void* O,
                    bool resid, bool fixed, double dt_step, double dt_pred, int slot, hipEvent_t e0, hipEvent_t e1)
{
    HfStepArgs<R> a = step_args<R>(c, item0, nitems, A, G, O);
    a.use_fixed = fixed;
    a.fixed_step = dt_step;
    a.fixed_pred = dt_pred;
    a.sync = c->d_sync + slot * HF_SYNC_WORDS;
    a.nmain = nmain;
    a.nrounds = kind == hf::HF_LAUNCH_MAIN ? 0 : c->nrounds;
    // the step's main launch skips the blocks the ghost fill counted
    if (slot == HF_SLOT_MAIN && !fixed) a.nskip = c->d_sync + HF_SLOT_MAIN * HF_SYNC_WORDS + HF_SYNC_NSKIP;
    const int grid = std::max(1, std::min(c->wg_cap, (nitems + 3) / 4));
    HF_HIP(hf::launch_update<R>(c->scheme, c->limiter, resid, kind, a, grid, c->stream, e0, e1));
    return HF_OK;
}

int launch_update(hf_ctx* c, int kind, int item0, int nitems, int nmain, const void* A, const void* G, void* O,
                  bool resid, bool fixed, double dt_step, double dt_pred, int slot, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr)
{
    if (c->prec == HF_FP64)
        return launch_update_r<double>(c, kind, item0, nitems, nmain, A, G, O, resid, fixed, dt_step, dt_pred, slot,
                                       e0, e1);
    return launch_update_r<float>(c, kind, item0, nitems, nmain, A, G, O, resid, fixed, dt_step, dt_pred, slot, e0,
                                  e1);
}

// A launch outside the device-resident loop (class-level API): its sync
// counters are cleared from the host first.
int launch_update_host(hf_ctx* c, int item0, int nitems, const void* A, const void* G, void* O, bool resid,
                       double dt_step, double dt_pred, bool timed)
{
    HF_HIP(hipMemsetAsync(c->d_sync + HF_SLOT_HOST * HF_SYNC_WORDS, 0, sizeof(uint32_t) * HF_SYNC_WORDS, c->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timed && nitems > 0) {
        const int rc = take_events(c, e0, e1);
        if (rc) return rc;
    }
    const int rc = launch_update(c, hf::HF_LAUNCH_MAIN, item0, nitems, nitems, A, G, O, resid, true, dt_step, dt_pred,
                                 HF_SLOT_HOST, e0, e1);
    if (e0) c->ev_pending.emplace_back(e0, e1);
    return rc;
}

int launch_ghost(hf_ctx* c, const HfGhost* tab, int n, const void* S, void* D, unsigned long long* gmax,
                 const GhostBook& book)
{
    if (c->prec == HF_FP64)
        HF_HIP(hf::launch_ghost<double>(tab, n, c->d_blocks, (const double*)S, (double*)D, (const double*)c->d_consts,
                                        gmax, c->gamma, book, c->stream));
    else
        HF_HIP(hf::launch_ghost<float>(tab, n, c->d_blocks, (const float*)S, (float*)D, (const float*)c->d_consts,
                                       gmax, c->gamma, book, c->stream));
    return HF_OK;
}

int launch_scan(hf_ctx* c, int b, const void* I, const void* Gh, int interior_only, unsigned long long* out)
{
    if (c->prec == HF_FP64)
        HF_HIP(hf::launch_scan<double>(c->blocks[b], (const double*)I, (const double*)Gh, interior_only, c->gamma, out,
                                       c->stream));
    else
        HF_HIP(hf::launch_scan<float>(c->blocks[b], (const float*)I, (const float*)Gh, interior_only, c->gamma, out,
                                      c->stream));
    return HF_OK;
}

GhostBook no_book()
{
    GhostBook bk;
    std::memset(&bk, 0, sizeof bk);
    return bk;
}

// Copy every ghost cell from S to D (same layout).
int copy_ghosts(hf_ctx* c, const void* S, void* D)
{
    return launch_ghost(c, c->d_copy, c->ncopy, S, D, c->d_scan + 2 * c->nb, no_book());
}

// Full TimeStep over problem p: interior and ghosts of block q come from
// src(q); returns the counted dt (timestep.cpp:29-68).
template <class F>
int full_dt(hf_ctx* c, int p, F src, double* dt)
{
    const HfProbDev& P = c->probs[p];
    HF_HIP(hipMemsetAsync(c->d_scan, 0, sizeof(unsigned long long) * 2 * c->nb, c->stream));
    for (int q = P.first; q < P.first + P.nb; q++) {
        const void* s = src(q);
        int rc = launch_scan(c, q, s, s, 0, c->d_scan + 2 * q);
        if (rc) return rc;
    }
    std::vector<unsigned long long> mx(2 * c->nb);
    HF_HIP(hipMemcpyAsync(mx.data(), c->d_scan, sizeof(unsigned long long) * 2 * c->nb, hipMemcpyDeviceToHost,
                          c->stream));
    HF_HIP(hipStreamSynchronize(c->stream));
    double dtx = 1e16, dty = 1e16;
    for (int q = P.first; q < P.first + P.nb; q++) {
        dtx = smin(dtx, c->blocks[q].dx / bitsd(mx[2 * q]));
        dty = smin(dty, c->blocks[q].dy / bitsd(mx[2 * q + 1]));
    }
    const int64_t k = ++c->calls[p];
    const double cfl_eff = (k <= 10) ? 0.2 * c->cfl : c->cfl;
    *dt = cfl_eff * smin(dtx, dty);
    if (p == 0 && k >= 1 && k <= c->hist_cap)
        HF_HIP(hipMemcpy(c->d_hist + (k - 1), dt, sizeof(double), hipMemcpyHostToDevice));
    return HF_OK;
}

// Make imax[cur] describe buf[cur]'s interior and zero gmax[cur].
int ensure_partials(hf_ctx* c)
{
    if (c->partials_valid) return HF_OK;
    HF_HIP(hipMemsetAsync(c->imax(c->cur), 0, sizeof(unsigned long long) * 2 * c->nb, c->stream));
    HF_HIP(hipMemsetAsync(c->gmax(c->cur), 0, sizeof(unsigned long long) * 2 * c->nb, c->stream));
    HF_HIP(hipMemsetAsync(c->d_imax_fix, 0, sizeof(unsigned long long) * 4 * c->nb, c->stream));
    HF_HIP(hipMemsetAsync(c->d_fixed, 0, sizeof(uint32_t) * 2 * c->nb, c->stream));
    HF_HIP(hipMemsetAsync(c->d_skip_next, 0, sizeof(uint32_t) * c->nb, c->stream));
    for (int b = 0; b < c->nb; b++) {
        int rc = launch_scan(c, b, c->buf[c->cur], c->buf[c->cur], 1, c->imax(c->cur) + 2 * b);
        if (rc) return rc;
    }
    c->partials_valid = true;
    return HF_OK;
}

// Queue nsteps device-resident steps starting from buffer parity c0.  With
// fixed, every step and MUSCL predictor uses dt and no TimeStep call is
// counted (no fix-up launches are needed).
int enqueue_steps(hf_ctx* c, int nsteps, bool timed, bool fixed = false, double dt = 0.0)
{
    const int c0 = c->cur;
    const bool timing = timed;
    for (int s = 0; s < nsteps; s++) {
        timed = timing && (c->prof_step++ % c->profile == 0);
        const int k = c->cur;
        void* A = c->buf[k];
        void* B = c->buf[k ^ 1];
        GhostBook bk;
        bk.calls = c->d_calls;
        bk.probs = c->d_probs;
        bk.nprob = c->np;
        bk.bump = fixed ? 0 : 1;
        bk.clear0 = c->gmax(k ^ 1);
        bk.clear1 = c->imax(k ^ 1);
        bk.nclear = 2 * c->nb;
        bk.queue = c->d_sync;
        bk.nqueue = HF_SLOTS * HF_SYNC_WORDS;
        bk.pending = c->d_pending;
        bk.npending = c->npending;
        bk.nblocks = c->nb;
        bk.imax_prev = c->imax(k);
        bk.imax_fix = c->d_imax_fix;
        bk.fixed = c->d_fixed;
        bk.skip = c->d_skip;
        bk.skip_next = c->d_skip_next;
        bk.fixstats = c->d_fixstats;
        int rc = launch_ghost(c, c->d_fill, c->nfill, A, B, c->gmax(k), bk);
        if (rc) return rc;
        // the main list, then with the CFL dt of a problem of several MUSCL
        // blocks the guess and check launches; timed from the first
        // launch's start to the last one's end
        const bool rounds = !fixed && c->nrounds > 0;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timed) {
            rc = take_events(c, e0, e1);
            if (rc) return rc;
        }
        const int n = c->main_items + c->fix_items;
        rc = launch_update(c, hf::HF_LAUNCH_MAIN, c->main_begin, c->main_items, c->main_items, A, B, B, false, fixed,
                           dt, dt, HF_SLOT_MAIN, e0, rounds ? nullptr : e1);
        if (rc) return rc;
        if (rounds) {
            rc = launch_update(c, hf::HF_LAUNCH_GUESS, c->main_begin, n, c->main_items, A, B, B, false, false, 0.0,
                               0.0, HF_SLOT_GUESS);
            if (rc) return rc;
            rc = launch_update(c, hf::HF_LAUNCH_CHECK, c->main_begin, n, c->main_items, A, B, B, false, false, 0.0,
                               0.0, HF_SLOT_CHECK, nullptr, e1);
            if (rc) return rc;
        }
        if (timed) c->ev_pending.emplace_back(e0, e1);
        c->cur ^= 1;
    }
    c->cur = c0;
    return HF_OK;
}

template <class R>
void to_device_layout(const hf_ctx* c, int b, const double* src, std::vector<R>& dst)
{
    const HfBlockDev& B = c->blocks[b];
    dst.assign((size_t)B.plane * 4, R(0));
    for (int v = 0; v < 4; v++)
        for (int y = 0; y < B.ty; y++)
            for (int x = 0; x < B.tx; x++)
                dst[(size_t)v * B.plane + (size_t)y * B.pitch + x] = (R)src[((size_t)v * B.ty + y) * B.tx + x];
}

template <class R>
void from_device_layout(const hf_ctx* c, int b, const std::vector<R>& src, double* dst)
{
    const HfBlockDev& B = c->blocks[b];
    for (int v = 0; v < 4; v++)
        for (int y = 0; y < B.ty; y++)
            for (int x = 0; x < B.tx; x++)
                dst[((size_t)v * B.ty + y) * B.tx + x] = (double)src[(size_t)v * B.plane + (size_t)y * B.pitch + x];
}

int download_block(hf_ctx* c, int b, const void* base, int64_t off, double* out)
{
    const HfBlockDev& B = c->blocks[b];
    HF_HIP(hipStreamSynchronize(c->stream));
    if (c->prec == HF_FP64) {
        std::vector<double> tmp((size_t)B.plane * 4);
        HF_HIP(hipMemcpy(tmp.data(), (const double*)base + off, tmp.size() * 8, hipMemcpyDeviceToHost));
        from_device_layout<double>(c, b, tmp, out);
    } else {
        std::vector<float> tmp((size_t)B.plane * 4);
        HF_HIP(hipMemcpy(tmp.data(), (const float*)base + off, tmp.size() * 4, hipMemcpyDeviceToHost));
        from_device_layout<float>(c, b, tmp, out);
    }
    return HF_OK;
}

void destroy_graphs(hf_ctx* c)
{
    for (auto& kv : c->graphs) {
        (void)hipGraphExecDestroy(kv.second.second);
        (void)hipGraphDestroy(kv.second.first);
    }
    c->graphs.clear();
}

// The device plan table of the update launches, item positions relative to
// the main list's first item: the first item of each fix-up round at
// [1..nrounds] and the end of the last round at [nrounds + 1]; then from
// segs_offset() the main list's (first, count) per phase and block, and from
// fixrange_offset() each block's (first, count) in the fix-up rounds.
std::vector<int32_t> plan_table(const hf_ctx& c)
{
    std::vector<int32_t> rt(c.nrounds + 2, 0);
    for (int l = 1; l <= c.nrounds; l++) rt[l] = c.fix_rounds[l - 1].first - c.main_begin;
    rt[c.nrounds + 1] = c.main_items + c.fix_items;
    rt.insert(rt.end(), c.segs.begin(), c.segs.end());
    rt.insert(rt.end(), c.fixrange.begin(), c.fixrange.end());
    return rt;
}

// The host-side plan of a case: block records and their layout in a state
// buffer, problems, the work items of the update launches (main launch and
// fix-up launches) and the resident-wave count they are sized for.  With
// dense, blocks are stored back to back without row padding, the layout of
// the arrays the C ABI exchanges.
int build_plan(hf_ctx* c, const hf_case_desc* desc, bool dense)
{
    c->scheme = desc->scheme;
    c->limiter = desc->limiter;
    c->prec = desc->precision;
    c->dt_mode = desc->dt_mode;
    c->gamma = desc->gamma;
    c->cfl = desc->cfl;
    c->g = c->scheme == HF_MUSCL_HANCOCK ? 2 : 1;
    c->nb = desc->nblocks;
    c->esize = c->prec == HF_FP64 ? 8 : 4;
    c->bdesc.assign(desc->blocks, desc->blocks + desc->nblocks);
    c->np = c->bdesc.back().problem + 1;

    // block layout: planes [var][ty][pitch], pitch a multiple of 32 elements
    // (dense: pitch = nx+2g and blocks back to back, the layout of the C ABI's arrays)
    int64_t off = 0;
    for (int b = 0; b < c->nb; b++) {
        const hf_block_desc& D = c->bdesc[b];
        HfBlockDev B;
        std::memset(&B, 0, sizeof B);
        B.nx = D.nx;
        B.ny = D.ny;
        B.g = c->g;
        B.tx = D.nx + 2 * c->g;
        B.ty = D.ny + 2 * c->g;
        B.pitch = dense ? B.tx : (int32_t)round_up(B.tx, 32);
        B.plane = (int64_t)B.pitch * B.ty;
        B.off = off;
        B.problem = D.problem;
        B.dx = (D.x_right - D.x_left) / D.nx;     // MeshBlock::calculate_dx (mesh_block.cpp:227-230)
        B.dy = (D.y_top - D.y_bottom) / D.ny;
        off = dense ? off + 4 * B.plane : round_up(off + 4 * B.plane, 64);
        c->blocks.push_back(B);
        c->resid_elems = std::max<int64_t>(c->resid_elems, 4 * B.plane);
    }
    c->total = off;
    // problems
    for (int b = 0; b < c->nb; b++) {
        const int p = c->blocks[b].problem;
        if ((int)c->probs.size() <= p) c->probs.push_back(HfProbDev{b, 0, 0, 0});
        c->blocks[b].lb = c->probs[p].nb++;
    }
    for (auto& P : c->probs) P.cps = c->scheme == HF_MUSCL_HANCOCK ? 1 + P.nb : 1;
    c->calls.assign(c->np, 0);

    // resident waves of a persistent update launch (the work items are sized for them)
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return fail(HF_EHIP, "hf_create: no usable GPU");
        const int per_cu = c->prec == HF_FP64 ? hf::update_wgs_per_cu<double>(c->scheme, c->limiter, false)
                                              : hf::update_wgs_per_cu<float>(c->scheme, c->limiter, false);
        c->wg_cap = std::max(1, cus * per_cu);
        // in-launch waits give up after 10 ms of the counter kernels read
        // with s_memrealtime (its rate in kHz)
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            khz = 100000;
        c->wait_ticks = (unsigned long long)khz * 10ull;
    }
    const int64_t wave_slots = 4LL * c->wg_cap;

    // work items: every block in order (the main launch of a step), then,
    // for compat MUSCL problems of several blocks, the lb-th blocks of all
    // problems for each fix-up launch lb = 1, 2, ..
    const int W = c->g == 2 ? 60 : 62;
    const int rows = desc->tile_rows > 0 ? desc->tile_rows : (c->scheme == HF_MUSCL_HANCOCK ? 48 : 32);
    for (auto& P : c->probs) c->maxnb = std::max(c->maxnb, P.nb);
    c->item_begin.assign(c->nb, 0);
    c->item_count.assign(c->nb, 0);
    // Items are taken from a queue in list order, so the list ends with small
    // items: the last round of a launch then finishes together instead of
    // waiting on a few long items (a timeline of the 8-block channel's main
    // launch showed half of the SIMDs idle over its last fifth with a 3/4
    // split into `rows`- and 32-row items).  Main launch: `rows`-row items,
    // then rows/2, then 8-row items (on the channel 11/16, 1/4 and 1/16 of the
    // rows, measured fastest among eight such plans; launches of many rounds
    // keep the small phases to about one round's worth of work).  A
    // fix-up launch computes a single block on the whole GPU: uniform items
    // sized so that one round of the resident waves takes them all.
    // A MUSCL block whose last column strip would carry at most TR_COLS of its
    // 60 output columns (1024 = 17 x 60 + 4) gets that band as transposed
    // items instead: lanes along rows, the wave walking the band's columns,
    // 60 rows per item (k_muscl's muscl_item_t; rows = -(cols + 256 * nrows)).
    constexpr int TR_COLS = 30;
    auto add_rows = [&](int b, int y0, int y1, int r, bool tr) {
        const HfBlockDev& B = c->blocks[b];
        int strips = (B.nx + W - 1) / W;
        const int rem = B.nx - (strips - 1) * W;
        const bool band = tr && c->scheme == HF_MUSCL_HANCOCK && strips > 1 && rem <= TR_COLS;
        if (band) strips--;
        for (int y = y0; y < y1; y += r)
            for (int s = 0; s < strips; s++) c->items.push_back(HfItem{b, s * W, c->g + y, std::min(r, y1 - y)});
        if (band)
            for (int y = y0; y < y1; y += 60)
                c->items.push_back(HfItem{b, c->g + strips * W, c->g + y - 2, -(rem + 256 * std::min(60, y1 - y))});
    };
    // column strips of a block without its transposed band (the band is cheap)
    auto full_strips = [&](int b) {
        const HfBlockDev& B = c->blocks[b];
        const int strips = (B.nx + W - 1) / W;
        return (c->scheme == HF_MUSCL_HANCOCK && strips > 1 && B.nx - (strips - 1) * W <= TR_COLS) ? strips - 1
                                                                                                   : strips;
    };
    for (int b = 0; b < c->nb; b++) {
        c->item_begin[b] = (int)c->items.size();
        add_rows(b, 0, c->blocks[b].ny, rows, false);   // contiguous per block (class-level launches)
        c->item_count[b] = (int)c->items.size() - c->item_begin[b];
    }
    const int perblock_end = (int)c->items.size();
    {
        // wave-rows of the launch; the last two phases hold about 17 and 4.25
        // rows per resident wave (the 8-block channel: 1/4 and 1/16 of its
        // rows), at most 1/4 and 1/16 of each block
        double work = 0.0;
        for (int b = 0; b < c->nb; b++) work += (double)full_strips(b) * c->blocks[b].ny;
        const double f3 = std::min(1.0 / 16, 4.25 * (double)wave_slots / std::max(1.0, work));
        const double f2 = std::min(0.25, 17.0 * (double)wave_slots / std::max(1.0, work));
        const int tail = desc->tail_rows > 0 ? desc->tail_rows : 8;
        const int r2 = std::max(1, std::min(rows, std::max(tail, rows / 2))), r3 = std::max(1, std::min(rows, tail));
        std::vector<int> y1(c->nb), y2(c->nb);
        for (int b = 0; b < c->nb; b++) {
            const int ny = c->blocks[b].ny;
            y1[b] = std::min(ny, (int)(ny * (1.0 - f2 - f3)) / rows * rows);
            y2[b] = std::min(ny, y1[b] + std::max(0, (int)(ny * (1.0 - f3)) - y1[b]) / r2 * r2);
        }
        c->segs.assign(6 * (size_t)c->nb, 0);
        for (int ph = 0; ph < 3; ph++)
            for (int b = 0; b < c->nb; b++) {
                const int first = (int)c->items.size();
                if (ph == 0) add_rows(b, 0, y1[b], rows, true);
                else if (ph == 1) add_rows(b, y1[b], y2[b], r2, true);
                else add_rows(b, y2[b], c->blocks[b].ny, r3, true);
                c->segs[2 * (ph * c->nb + b)] = first - perblock_end;
                c->segs[2 * (ph * c->nb + b) + 1] = (int)c->items.size() - first;
            }
    }
    c->main_begin = perblock_end;
    c->main_items = (int)c->items.size() - perblock_end;
    // Wave maxima read their slot before raising it where hundreds of a
    // block's items end together (raise_max in hf_kernels.h); not for
    // Godunov, whose items are short, nor for blocks of few items.
    int most = 0;
    for (int b = 0; b < c->nb; b++) most = std::max(most, c->item_count[b]);
    c->probe_max = (c->scheme == HF_MUSCL_HANCOCK && most >= 256) ? 1 : 0;
    c->fixrange.assign(2 * (size_t)c->nb, 0);
    if (c->scheme == HF_MUSCL_HANCOCK && c->dt_mode == HF_DT_COMPAT)
        for (int lb = 1; lb < c->maxnb; lb++) {
            // rows per item so that the launch's wave-rows fill about 95 % of
            // one round of the resident waves, within 4..16 rows
            int64_t work = 0;
            for (const auto& P : c->probs)
                if (P.nb > lb) {
                    work += (int64_t)full_strips(P.first + lb) * c->blocks[P.first + lb].ny;
                }
            const int64_t target = std::max<int64_t>(1, wave_slots * 95 / 100);
            const int fr = desc->fix_rows > 0
                               ? desc->fix_rows
                               : (int)std::max<int64_t>(4, std::min<int64_t>(16, (work + target - 1) / target));
            const int first = (int)c->items.size();
            for (const auto& P : c->probs)
                if (P.nb > lb) {
                    const int b = P.first + lb, bfirst = (int)c->items.size();
                    add_rows(b, 0, c->blocks[b].ny, std::min(rows, fr), true);
                    c->fixrange[2 * b] = bfirst - c->main_begin;
                    c->fixrange[2 * b + 1] = (int)c->items.size() - bfirst;
                }
            c->fix_rounds.emplace_back(first, (int)c->items.size() - first);
        }
    // the fix-up rounds follow the main list in the item array
    c->nrounds = (int)c->fix_rounds.size();
    c->fix_items = (int)c->items.size() - (c->main_begin + c->main_items);
    c->npending = std::max(1, (c->fix_items + 31) / 32);

    return HF_OK;
}

}  // namespace

// =================================================================== C ABI ==

extern "C" {

const char* hf_last_error(void) { return g_err.c_str(); }

const char* hf_version(void) { return "hyperflow_amd 0.1 gfx950"; }

int hf_create(const hf_case_desc* desc, hf_ctx** out)
{
    if (!out) return fail(HF_EINVAL, "hf_create: null output pointer");
    *out = nullptr;
    int rc = validate(desc);
    if (rc) return rc;
    std::unique_ptr<hf_ctx> c(new hf_ctx());
    rc = build_plan(c.get(), desc, false);
    if (rc) return rc;

    std::vector<HfGhost> fill, copy;
    build_tables(*c, fill, copy);
    c->nfill = (int)fill.size();
    c->ncopy = (int)copy.size();

    hf_ctx* C = c.get();
    auto cleanup_fail = [&](int code) {
        hf_destroy(c.release());
        return code;
    };
#define HF_HIPC(call)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            fail(HF_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));                 \
            return cleanup_fail(HF_EHIP);                                                     \
        }                                                                                     \
    } while (0)
    HF_HIPC(hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking));
    for (int i = 0; i < 2; i++) {
        HF_HIPC(hipMalloc(&C->buf[i], (size_t)C->total * C->esize));
        HF_HIPC(hipMemset(C->buf[i], 0, (size_t)C->total * C->esize));
    }
    HF_HIPC(hipMalloc(&C->d_blocks, sizeof(HfBlockDev) * C->nb));
    HF_HIPC(hipMemcpy(C->d_blocks, C->blocks.data(), sizeof(HfBlockDev) * C->nb, hipMemcpyHostToDevice));
    HF_HIPC(hipMalloc(&C->d_probs, sizeof(HfProbDev) * C->np));
    HF_HIPC(hipMemcpy(C->d_probs, C->probs.data(), sizeof(HfProbDev) * C->np, hipMemcpyHostToDevice));
    HF_HIPC(hipMalloc(&C->d_items, sizeof(HfItem) * C->items.size()));
    HF_HIPC(hipMemcpy(C->d_items, C->items.data(), sizeof(HfItem) * C->items.size(), hipMemcpyHostToDevice));
    HF_HIPC(hipMalloc(&C->d_fill, sizeof(HfGhost) * fill.size()));
    HF_HIPC(hipMemcpy(C->d_fill, fill.data(), sizeof(HfGhost) * fill.size(), hipMemcpyHostToDevice));
    HF_HIPC(hipMalloc(&C->d_copy, sizeof(HfGhost) * copy.size()));
    HF_HIPC(hipMemcpy(C->d_copy, copy.data(), sizeof(HfGhost) * copy.size(), hipMemcpyHostToDevice));
    {
        std::vector<double> cd(16 * (size_t)C->nb);
        for (int b = 0; b < C->nb; b++)
            for (int s = 0; s < 4; s++)
                for (int v = 0; v < 4; v++) cd[(b * 4 + s) * 4 + v] = C->bdesc[b].state[s][v];
        if (C->prec == HF_FP64) {
            HF_HIPC(hipMalloc(&C->d_consts, cd.size() * 8));
            HF_HIPC(hipMemcpy(C->d_consts, cd.data(), cd.size() * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> cf(cd.begin(), cd.end());
            HF_HIPC(hipMalloc(&C->d_consts, cf.size() * 4));
            HF_HIPC(hipMemcpy(C->d_consts, cf.data(), cf.size() * 4, hipMemcpyHostToDevice));
        }
    }
    HF_HIPC(hipMalloc(&C->d_imax, sizeof(unsigned long long) * 4 * C->nb));
    HF_HIPC(hipMalloc(&C->d_gmax, sizeof(unsigned long long) * 4 * C->nb));
    HF_HIPC(hipMalloc(&C->d_scan, sizeof(unsigned long long) * 4 * C->nb));
    HF_HIPC(hipMemset(C->d_imax, 0, sizeof(unsigned long long) * 4 * C->nb));
    HF_HIPC(hipMemset(C->d_gmax, 0, sizeof(unsigned long long) * 4 * C->nb));
    HF_HIPC(hipMalloc(&C->d_imax_fix, sizeof(unsigned long long) * 4 * C->nb));
    HF_HIPC(hipMemset(C->d_imax_fix, 0, sizeof(unsigned long long) * 4 * C->nb));
    HF_HIPC(hipMalloc(&C->d_fixed, sizeof(uint32_t) * 2 * C->nb));
    HF_HIPC(hipMalloc(&C->d_fixstats, sizeof(unsigned long long) * 4));
    HF_HIPC(hipMemset(C->d_fixstats, 0, sizeof(unsigned long long) * 4));
    HF_HIPC(hipMemset(C->d_fixed, 0, sizeof(uint32_t) * 2 * C->nb));
    for (uint32_t** p : {&C->d_skip, &C->d_skip_next}) {
        HF_HIPC(hipMalloc(p, sizeof(uint32_t) * C->nb));
        HF_HIPC(hipMemset(*p, 0, sizeof(uint32_t) * C->nb));
    }
    HF_HIPC(hipMalloc(&C->d_sync, sizeof(uint32_t) * HF_SLOTS * HF_SYNC_WORDS));
    HF_HIPC(hipMemset(C->d_sync, 0, sizeof(uint32_t) * HF_SLOTS * HF_SYNC_WORDS));
    {
        const std::vector<int32_t> rt = plan_table(*C);
        HF_HIPC(hipMalloc(&C->d_rounds, sizeof(int32_t) * rt.size()));
        HF_HIPC(hipMemcpy(C->d_rounds, rt.data(), sizeof(int32_t) * rt.size(), hipMemcpyHostToDevice));
    }
    HF_HIPC(hipMalloc(&C->d_pending, sizeof(uint32_t) * C->npending));
    HF_HIPC(hipMemset(C->d_pending, 0, sizeof(uint32_t) * C->npending));
    HF_HIPC(hipMalloc(&C->d_calls, sizeof(uint32_t) * C->np));
    HF_HIPC(hipMemset(C->d_calls, 0, sizeof(uint32_t) * C->np));
    HF_HIPC(hipHostMalloc((void**)&C->h_calls_pinned, sizeof(uint32_t) * C->np, hipHostMallocDefault));
    HF_HIPC(hipEventCreateWithFlags(&C->calls_copied, hipEventDisableTiming));
    HF_HIPC(hipEventRecord(C->calls_copied, C->stream));
    HF_HIPC(hipMalloc(&C->d_hist, sizeof(double) * C->hist_cap));
    HF_HIPC(hipMemset(C->d_hist, 0, sizeof(double) * C->hist_cap));
    HF_HIPC(hipMalloc(&C->d_part, sizeof(double) * 8 * C->part_wg));
    HF_HIPC(hipMalloc(&C->d_resid, (size_t)C->resid_elems * C->esize));
#undef HF_HIPC
    *out = c.release();
    return HF_OK;
}

void hf_destroy(hf_ctx* c)
{
    if (!c) return;
    (void)g_exit_hook;
    if (hf_process_exiting()) {   // see hf_process_exiting: host memory only
        delete c;
        return;
    }
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    destroy_graphs(c);
    for (auto& e : c->ev_pending) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    for (auto& e : c->ev_free) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    for (void* p : {c->buf[0], c->buf[1], (void*)c->d_blocks, (void*)c->d_probs, (void*)c->d_items, (void*)c->d_fill,
                    (void*)c->d_copy, c->d_consts, (void*)c->d_imax, (void*)c->d_gmax, (void*)c->d_scan,
                    (void*)c->d_calls, (void*)c->d_hist, (void*)c->d_part, c->d_resid, (void*)c->d_imax_fix,
                    (void*)c->d_fixed, (void*)c->d_fixstats, (void*)c->d_sync, (void*)c->d_rounds, (void*)c->d_pending, (void*)c->d_skip,
                    (void*)c->d_skip_next})
        if (p) (void)hipFree(p);
    if (c->h_calls_pinned) (void)hipHostFree(c->h_calls_pinned);
    if (c->calls_copied) (void)hipEventDestroy(c->calls_copied);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int hf_ghost_cells(const hf_ctx* c)
{
    if (!c) return fail(HF_EINVAL, "null context");
    return c->g;
}

int hf_upload(hf_ctx* c, int b, const double* fields, size_t n)
{
    if (!c || !fields) return fail(HF_EINVAL, "hf_upload: null argument");
    if (b < 0 || b >= c->nb) return fail(HF_EINVAL, "hf_upload: block out of range");
    const HfBlockDev& B = c->blocks[b];
    if (n != (size_t)4 * B.tx * B.ty) return fail(HF_EINVAL, "hf_upload: size must be 4*(nx+2g)*(ny+2g)");
    HF_HIP(hipStreamSynchronize(c->stream));
    if (c->prec == HF_FP64) {
        std::vector<double> tmp;
        to_device_layout<double>(c, b, fields, tmp);
        HF_HIP(hipMemcpy((double*)c->buf[c->cur] + B.off, tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp;
        to_device_layout<float>(c, b, fields, tmp);
        HF_HIP(hipMemcpy((float*)c->buf[c->cur] + B.off, tmp.data(), tmp.size() * 4, hipMemcpyHostToDevice));
    }
    c->ghosts_pending = false;
    c->partials_valid = false;
    return HF_OK;
}

int hf_download(hf_ctx* c, int b, double* fields, size_t n)
{
    if (!c || !fields) return fail(HF_EINVAL, "hf_download: null argument");
    if (b < 0 || b >= c->nb) return fail(HF_EINVAL, "hf_download: block out of range");
    const HfBlockDev& B = c->blocks[b];
    if (n != (size_t)4 * B.tx * B.ty) return fail(HF_EINVAL, "hf_download: size must be 4*(nx+2g)*(ny+2g)");
    return download_block(c, b, c->buf[c->cur], B.off, fields);
}

int hf_apply_bcs(hf_ctx* c)
{
    if (!c) return fail(HF_EINVAL, "null context");
    void* A = c->buf[c->cur];
    void* B = c->buf[c->cur ^ 1];
    HF_HIP(hipMemsetAsync(c->d_scan, 0, sizeof(unsigned long long) * 4 * c->nb, c->stream));
    int rc = launch_ghost(c, c->d_fill, c->nfill, A, B, c->d_scan, no_book());
    if (rc) return rc;
    rc = copy_ghosts(c, B, A);
    if (rc) return rc;
    c->ghosts_pending = true;
    c->partials_valid = false;
    return HF_OK;
}

int hf_time_step(hf_ctx* c, int p, double* dt)
{
    if (!c || !dt) return fail(HF_EINVAL, "hf_time_step: null argument");
    if (p < 0 || p >= c->np) return fail(HF_EINVAL, "hf_time_step: problem out of range");
    void* A = c->buf[c->cur];
    return full_dt(c, p, [&](int) { return A; }, dt);
}

int hf_scheme_residual(hf_ctx* c, int b, double* L, size_t n)
{
    if (!c || !L) return fail(HF_EINVAL, "hf_scheme_residual: null argument");
    if (b < 0 || b >= c->nb) return fail(HF_EINVAL, "hf_scheme_residual: block out of range");
    const HfBlockDev& B = c->blocks[b];
    if (n != (size_t)4 * B.tx * B.ty) return fail(HF_EINVAL, "hf_scheme_residual: size must be 4*(nx+2g)*(ny+2g)");
    void* A = c->buf[c->cur];
    double dt_pred = 0.0;
    if (c->scheme == HF_MUSCL_HANCOCK) {
        int rc = full_dt(c, B.problem, [&](int) { return A; }, &dt_pred);   // muscl_hancock.cpp:246
        if (rc) return rc;
    }
    HF_HIP(hipMemsetAsync(c->d_resid, 0, (size_t)c->resid_elems * c->esize, c->stream));
    int rc = launch_update_host(c, c->item_begin[b], c->item_count[b], A, A, c->d_resid, true, 0.0, dt_pred, false);
    if (rc) return rc;
    return download_block(c, b, c->d_resid, 0, L);
}

int hf_advance(hf_ctx* c, const double* dt)
{
    if (!c || !dt) return fail(HF_EINVAL, "hf_advance: null argument");
    void* A = c->buf[c->cur];
    void* B = c->buf[c->cur ^ 1];
    if (!c->ghosts_pending) {
        int rc = copy_ghosts(c, A, B);
        if (rc) return rc;
    }
    for (int p = 0; p < c->np; p++) {
        const HfProbDev& P = c->probs[p];
        for (int q = P.first; q < P.first + P.nb; q++) {
            double dt_pred = dt[p];
            if (c->scheme == HF_MUSCL_HANCOCK) {
                if (c->dt_mode == HF_DT_COMPAT) {
                    // the predictor's TimeStep sees blocks before q already updated
                    int rc = full_dt(c, p, [&](int r) { return r < q ? B : A; }, &dt_pred);
                    if (rc) return rc;
                } else {
                    const int64_t k = ++c->calls[p];
                    if (p == 0 && k <= c->hist_cap)
                        HF_HIP(hipMemcpy(c->d_hist + (k - 1), &dt_pred, sizeof(double), hipMemcpyHostToDevice));
                }
            }
            int rc = launch_update_host(c, c->item_begin[q], c->item_count[q], A, B, B, false, dt[p], dt_pred,
                                        c->profile != 0);
            if (rc) return rc;
        }
    }
    c->cur ^= 1;
    c->ghosts_pending = false;
    c->partials_valid = false;
    return HF_OK;
}

static int run_loop(hf_ctx* c, int nsteps, int flags, bool fixed, double dt)
{
    // fixed-dt steps never read the wave-speed maxima
    int rc = fixed ? HF_OK : ensure_partials(c);
    if (rc) return rc;
    HF_HIP(hipEventSynchronize(c->calls_copied));   // earlier upload must have consumed the staging buffer
    for (int p = 0; p < c->np; p++) c->h_calls_pinned[p] = (uint32_t)c->calls[p];
    HF_HIP(hipMemcpyAsync(c->d_calls, c->h_calls_pinned, sizeof(uint32_t) * c->np, hipMemcpyHostToDevice, c->stream));
    HF_HIP(hipEventRecord(c->calls_copied, c->stream));
    if (flags & HF_RUN_GRAPH) {
        uint64_t dtbits = 0;
        if (fixed) std::memcpy(&dtbits, &dt, sizeof dtbits);
        const auto key = std::make_tuple(nsteps, c->cur, dtbits);
        auto it = c->graphs.find(key);
        if (it == c->graphs.end()) {
            hipGraph_t g = nullptr;
            hipGraphExec_t ge = nullptr;
            HF_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_steps(c, nsteps, false, fixed, dt);
            hipError_t ce = hipStreamEndCapture(c->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            HF_HIP(ce);
            HF_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            it = c->graphs.emplace(key, std::make_pair(g, ge)).first;
        }
        HF_HIP(hipGraphLaunch(it->second.second, c->stream));
    } else {
        rc = enqueue_steps(c, nsteps, c->profile != 0, fixed, dt);
        if (rc) return rc;
    }
    c->cur ^= (nsteps & 1);
    if (!fixed)
        for (int p = 0; p < c->np; p++) c->calls[p] += (int64_t)nsteps * c->probs[p].cps;
    else
        c->partials_valid = false;
    c->ghosts_pending = false;
    return HF_OK;
}

int hf_run(hf_ctx* c, int nsteps, int flags)
{
    if (!c) return fail(HF_EINVAL, "null context");
    if (nsteps < 0) return fail(HF_EINVAL, "hf_run: nsteps must be >= 0");
    if (nsteps == 0) return HF_OK;
    return run_loop(c, nsteps, flags, false, 0.0);
}

int hf_run_fixed(hf_ctx* c, int nsteps, double dt, int flags)
{
    if (!c) return fail(HF_EINVAL, "null context");
    if (nsteps < 0) return fail(HF_EINVAL, "hf_run_fixed: nsteps must be >= 0");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(HF_EINVAL, "hf_run_fixed: dt must be finite and > 0");
    if (nsteps == 0) return HF_OK;
    return run_loop(c, nsteps, flags, true, dt);
}

int hf_synchronize(hf_ctx* c)
{
    if (!c) return fail(HF_EINVAL, "null context");
    HF_HIP(hipStreamSynchronize(c->stream));
    return HF_OK;
}

int hf_dt_history(hf_ctx* c, double* out, int max, int* count)
{
    if (!c) return fail(HF_EINVAL, "null context");
    const int64_t n = c->calls[0];
    if (count) *count = (int)std::min<int64_t>(n, INT32_MAX);
    if (out && max > 0) {
        HF_HIP(hipStreamSynchronize(c->stream));
        const int64_t k = std::min<int64_t>(std::min<int64_t>(n, max), c->hist_cap);
        if (k > 0) HF_HIP(hipMemcpy(out, c->d_hist, sizeof(double) * k, hipMemcpyDeviceToHost));
    }
    return HF_OK;
}

int hf_calls(hf_ctx* c, int p, int64_t* calls)
{
    if (!c || !calls) return fail(HF_EINVAL, "hf_calls: null argument");
    if (p < 0 || p >= c->np) return fail(HF_EINVAL, "hf_calls: problem out of range");
    *calls = c->calls[p];
    return HF_OK;
}

int hf_profile(hf_ctx* c, int enable)
{
    if (!c) return fail(HF_EINVAL, "null context");
    if (enable < 0) return fail(HF_EINVAL, "hf_profile: period must be >= 0");
    c->profile = enable;
    c->prof_step = 0;
    return HF_OK;
}

int hf_kernel_stats(hf_ctx* c, int64_t* launches, double* total_ms)
{
    if (!c) return fail(HF_EINVAL, "null context");
    HF_HIP(hipStreamSynchronize(c->stream));
    for (auto& e : c->ev_pending) {
        float ms = 0.f;
        HF_HIP(hipEventElapsedTime(&ms, e.first, e.second));
        c->prof_ms += ms;
        c->prof_launches++;
        c->ev_free.push_back(e);
    }
    c->ev_pending.clear();
    if (launches) *launches = c->prof_launches;
    if (total_ms) *total_ms = c->prof_ms;
    c->prof_launches = 0;
    c->prof_ms = 0.0;
    return HF_OK;
}

int hf_fixup_stats(hf_ctx* c, int64_t out[4])
{
    if (!c || !out) return fail(HF_EINVAL, "hf_fixup_stats: null argument");
    unsigned long long h[4] = {0, 0, 0, 0};
    if (c->d_fixstats) {
        HF_HIP(hipStreamSynchronize(c->stream));
        HF_HIP(hipMemcpy(h, c->d_fixstats, sizeof h, hipMemcpyDeviceToHost));
        HF_HIP(hipMemset(c->d_fixstats, 0, sizeof h));
    }
    for (int k = 0; k < 4; k++) out[k] = (int64_t)h[k];
    return HF_OK;
}

int hf_conservation(hf_ctx* c, int p, double sums[4])
{
    if (!c || !sums) return fail(HF_EINVAL, "hf_conservation: null argument");
    if (p < 0 || p >= c->np) return fail(HF_EINVAL, "hf_conservation: problem out of range");
    double S[4] = {0, 0, 0, 0}, Cm[4] = {0, 0, 0, 0};
    std::vector<double> part(8 * (size_t)c->part_wg);
    const HfProbDev& P = c->probs[p];
    for (int q = P.first; q < P.first + P.nb; q++) {
        if (c->prec == HF_FP64)
            HF_HIP(hf::launch_conserve<double>(c->blocks[q], (const double*)c->buf[c->cur], c->d_part, c->part_wg,
                                               c->stream));
        else
            HF_HIP(hf::launch_conserve<float>(c->blocks[q], (const float*)c->buf[c->cur], c->d_part, c->part_wg,
                                              c->stream));
        HF_HIP(hipMemcpyAsync(part.data(), c->d_part, part.size() * 8, hipMemcpyDeviceToHost, c->stream));
        HF_HIP(hipStreamSynchronize(c->stream));
        for (int w = 0; w < c->part_wg; w++)
            for (int k = 0; k < 4; k++) {
                for (int h = 0; h < 2; h++) {
                    const double v = part[(w * 4 + k) * 2 + h];
                    const double t = S[k] + v;
                    Cm[k] += (std::fabs(S[k]) >= std::fabs(v)) ? ((S[k] - t) + v) : ((v - t) + S[k]);
                    S[k] = t;
                }
            }
    }
    for (int k = 0; k < 4; k++) sums[k] = S[k] + Cm[k];
    return HF_OK;
}

int hf_step_noalloc(const hf_case_desc* desc, const void* d_in, void* d_out, int64_t* calls, double* d_dt,
                    void* stream)
{
    if (!desc || !d_in || !d_out || !calls) return fail(HF_EINVAL, "hf_step_noalloc: null argument");
    if (d_in == d_out) return fail(HF_EINVAL, "hf_step_noalloc: input and output must be different arrays");
    int rc = validate(desc);
    if (rc) return rc;
    // the host-side plan only: this context owns no device memory
    hf_ctx c;
    rc = build_plan(&c, desc, true);
    if (rc) return rc;
    std::vector<HfGhost> fill, copy;
    build_tables(c, fill, copy);
    for (int p = 0; p < c.np; p++)
        if (calls[p] < 0 || calls[p] > (int64_t)UINT32_MAX - c.probs[p].cps)
            return fail(HF_EINVAL, "hf_step_noalloc: TimeStep call counter out of range");

    // Scratch layout: every region 256-byte aligned.  The zeroed arrays and the
    // call counters are part of the host image, so one copy sets everything.
    size_t used = 0;
    auto take = [&](size_t bytes) {
        const size_t o = (size_t)round_up((int64_t)used, 256);
        used = o + bytes;
        return o;
    };
    const size_t nb = (size_t)c.nb, np = (size_t)c.np;
    const size_t o_blocks = take(sizeof(HfBlockDev) * nb), o_probs = take(sizeof(HfProbDev) * np),
                 o_items = take(sizeof(HfItem) * c.items.size()), o_fill = take(sizeof(HfGhost) * fill.size()),
                 o_consts = take(c.esize * 16 * nb), o_imax = take(sizeof(unsigned long long) * 4 * nb),
                 o_gmax = take(sizeof(unsigned long long) * 4 * nb),
                 o_fix = take(sizeof(unsigned long long) * 4 * nb), o_fixed = take(sizeof(uint32_t) * 2 * nb),
                 o_skip = take(sizeof(uint32_t) * nb), o_skipn = take(sizeof(uint32_t) * nb),
                 o_sync = take(sizeof(uint32_t) * HF_SLOTS * HF_SYNC_WORDS),
                 o_rounds = take(sizeof(int32_t) * plan_table(c).size()), o_pend = take(sizeof(uint32_t) * c.npending),
                 o_calls = take(sizeof(uint32_t) * np);
    if (used > HF_NOALLOC_SCRATCH_BYTES)
        return fail(HF_EINVAL, "hf_step_noalloc: the case needs " + std::to_string(used) +
                                   " bytes of device scratch, more than the maximum of " +
                                   std::to_string((size_t)HF_NOALLOC_SCRATCH_BYTES));
    void* base = nullptr;
    std::string err;
    if (hf_noalloc_scratch(&base, &err)) return fail(HF_EHIP, "hf_step_noalloc: " + err);

    std::vector<unsigned char> img(used, 0);
    std::memcpy(img.data() + o_blocks, c.blocks.data(), sizeof(HfBlockDev) * nb);
    std::memcpy(img.data() + o_probs, c.probs.data(), sizeof(HfProbDev) * np);
    std::memcpy(img.data() + o_items, c.items.data(), sizeof(HfItem) * c.items.size());
    std::memcpy(img.data() + o_fill, fill.data(), sizeof(HfGhost) * fill.size());
    for (size_t b = 0; b < nb; b++)
        for (int sd = 0; sd < 4; sd++)
            for (int v = 0; v < 4; v++) {
                const size_t k = (b * 4 + sd) * 4 + v;
                if (c.prec == HF_FP64) {
                    const double x = c.bdesc[b].state[sd][v];
                    std::memcpy(img.data() + o_consts + 8 * k, &x, 8);
                } else {
                    const float x = (float)c.bdesc[b].state[sd][v];
                    std::memcpy(img.data() + o_consts + 4 * k, &x, 4);
                }
            }
    for (size_t p = 0; p < np; p++) {
        const uint32_t k = (uint32_t)calls[p];
        std::memcpy(img.data() + o_calls + 4 * p, &k, 4);
    }
    {
        const std::vector<int32_t> rt = plan_table(c);
        std::memcpy(img.data() + o_rounds, rt.data(), sizeof(int32_t) * rt.size());
    }

    unsigned char* S = (unsigned char*)base;
    c.stream = (hipStream_t)stream;
    c.buf[0] = const_cast<void*>(d_in);   // read only: the ghost fill and the update write buf[1]
    c.buf[1] = d_out;
    c.cur = 0;
    c.d_blocks = (HfBlockDev*)(S + o_blocks);
    c.d_probs = (HfProbDev*)(S + o_probs);
    c.d_items = (HfItem*)(S + o_items);
    c.d_fill = (HfGhost*)(S + o_fill);
    c.nfill = (int)fill.size();
    c.d_consts = S + o_consts;
    c.d_imax = (unsigned long long*)(S + o_imax);
    c.d_gmax = (unsigned long long*)(S + o_gmax);
    c.d_imax_fix = (unsigned long long*)(S + o_fix);
    c.d_fixed = (uint32_t*)(S + o_fixed);
    c.d_skip = (uint32_t*)(S + o_skip);
    c.d_skip_next = (uint32_t*)(S + o_skipn);
    c.d_sync = (uint32_t*)(S + o_sync);
    c.d_rounds = (int32_t*)(S + o_rounds);
    c.d_pending = (uint32_t*)(S + o_pend);
    c.d_calls = (uint32_t*)(S + o_calls);
    c.d_hist = d_dt;
    c.hist_cap = d_dt ? c.probs[0].cps : 0;
    c.hist_base = calls[0];

    HF_HIP(hipMemcpyAsync(base, img.data(), used, hipMemcpyHostToDevice, c.stream));
    // wave-speed maxima of the input's interior (a device-resident loop gets
    // them from the previous step's update epilogue)
    for (int b = 0; b < c.nb; b++) {
        rc = launch_scan(&c, b, d_in, d_in, 1, c.imax(0) + 2 * b);
        if (rc) return rc;
    }
    rc = enqueue_steps(&c, 1, false);
    if (rc) return rc;
    // the host image must outlive its copy, and the caller gets a finished step
    HF_HIP(hipStreamSynchronize(c.stream));
    for (size_t p = 0; p < np; p++) calls[p] += c.probs[p].cps;
    return HF_OK;
}

}  // extern "C"
