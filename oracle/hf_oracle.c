/*
 * hf_oracle.c -- CPU oracle (FP64) for HyperFlow's block-Cartesian Euler step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every formula is written in the reference's evaluation order (binary64,
 * no contraction: build with -ffp-contract=off).  Each function cites the
 * reference file:line it restates; paths are relative to the reference
 * checkout, under src/fvm_2d_cart/ unless stated otherwise.
 */
#define _GNU_SOURCE   /* sincos */
#include "hf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846   /* the value of glibc's M_PI */
#endif
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int nx, ny, tx, ty;          /* interior and padded cell counts */
    double xl, xr, yb, yt, dx, dy;
    int kind[4], nbr[4];         /* sides: bottom, right, top, left */
    double state[4][4];
    double* U;                   /* [4][ty][tx] */
} blk_t;

struct hfo_mesh {
    int nb, g, scheme, limiter;
    double gamma, cfl;
    unsigned steps;              /* TimeStep::steps, starts at 1 (timestep.cpp:20) */
    blk_t* b;
    double* log;
    int nlog, caplog;
};

static int g_threads = 1;
void hfo_set_threads(int n) { g_threads = n > 0 ? n : 1; }

/* std::min / std::max semantics (NaN in the second argument is ignored by
 * min, NaN in the first is kept), as used at timestep.cpp:51-52,
 * riemann_hllc.cpp:146-147, limiter.cpp:58, minbee.cpp:37, vanleer.cpp:35. */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

#define AT(B, v, i, j) ((B)->U[((size_t)(v) * (B)->ty + (size_t)(j)) * (B)->tx + (size_t)(i)])

/* ---------------------------------------------------------------- model -- */

/* EulerIdealGasModel::cons_to_prim, model/euler_ideal_gas/euler_ideal_gas.cpp:76-92 */
void hfo_cons_to_prim(double gamma, const double* U, double* W)
{
    double rho = U[0], u = U[1] / rho, v = U[2] / rho;
    W[0] = rho;
    W[1] = u;
    W[2] = v;
    W[3] = (gamma - 1.0) * (U[3] - 0.5 * rho * (u * u + v * v));
}

/* EulerIdealGasModel::prim_to_cons, euler_ideal_gas.cpp:96-109 */
void hfo_prim_to_cons(double gamma, const double* W, double* U)
{
    double rho = W[0], u = W[1], v = W[2], p = W[3];
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v);
}

/* prim_total_energy with prim_internal_energy, euler_ideal_gas.cpp:32-46 */
static inline double prim_total_energy(double gamma, const double* W)
{
    double e = W[3] / (W[0] * (gamma - 1.0));
    return W[0] * (0.5 * (W[1] * W[1] + W[2] * W[2]) + e);
}

/* cons_flux_x / cons_flux_y, euler_ideal_gas.cpp:113-149: the energy is
 * re-derived from the primitives. */
void hfo_cons_flux(double gamma, const double* U, int dir, double* F)
{
    double W[4];
    hfo_cons_to_prim(gamma, U, W);
    double E = prim_total_energy(gamma, W);
    double rho = W[0], u = W[1], v = W[2], p = W[3];
    if (dir == 0) {
        F[0] = rho * u;
        F[1] = rho * u * u + p;
        F[2] = rho * u * v;
        F[3] = u * (E + p);
    } else {
        F[0] = rho * v;
        F[1] = rho * u * v;
        F[2] = rho * v * v + p;
        F[3] = v * (E + p);
    }
}

/* ------------------------------------------------------------------ HLLC -- */

/* HLLCEulerRiemannSolver::calc_U_k, scheme/riemann/riemann_hllc/riemann_hllc.cpp:37-63 */
static void calc_U_k(double gamma, const double* wK, double S_K, double S_star, int dir, double* UK)
{
    double dK = wK[0], uK = wK[1], vK = wK[2], pK = wK[3];
    double EK = prim_total_energy(gamma, wK);
    if (dir == 0) {
        double coef = dK * ((S_K - uK) / (S_K - S_star));
        UK[0] = coef;
        UK[1] = coef * S_star;
        UK[2] = coef * vK;
        UK[3] = coef * (EK / dK + (S_star - uK) * (S_star + (pK / (dK * (S_K - uK)))));
    } else {
        double coef = dK * ((S_K - vK) / (S_K - S_star));
        UK[0] = coef;
        UK[1] = coef * uK;
        UK[2] = coef * S_star;
        UK[3] = coef * (EK / dK + (S_star - vK) * (S_star + (pK / (dK * (S_K - vK)))));
    }
}

/* HLLCEulerRiemannSolver::calc_hllc_flux, riemann_hllc.cpp:169-212, with
 * calc_wave_speeds_direct_min_max (:127-150), calc_S_star (:67-97) and
 * calc_F_star (:153-164).  Inconsistent / NaN states give a zero flux. */
void hfo_hllc(double gamma, const double* UL, const double* UR, int dir, double* F)
{
    double wL[4], wR[4], FL[4], FR[4], ULs[4], URs[4];
    hfo_cons_to_prim(gamma, UL, wL);
    hfo_cons_to_prim(gamma, UR, wR);

    double a_L = sqrt(gamma * wL[3] / wL[0]);
    double a_R = sqrt(gamma * wR[3] / wR[0]);
    double velL = dir == 0 ? wL[1] : wL[2];
    double velR = dir == 0 ? wR[1] : wR[2];
    double S_L = smin(velL - a_L, velR - a_R);
    double S_R = smax(velL + a_L, velR + a_R);

    hfo_cons_flux(gamma, UL, dir, FL);
    hfo_cons_flux(gamma, UR, dir, FR);

    double dL = wL[0], pL = wL[3], dR = wR[0], pR = wR[3];
    double num = pR - pL + dL * velL * (S_L - velL) - dR * velR * (S_R - velR);
    double denom = dL * (S_L - velL) - dR * (S_R - velR);
    double S_star = num / denom;

    calc_U_k(gamma, wL, S_L, S_star, dir, ULs);
    calc_U_k(gamma, wR, S_R, S_star, dir, URs);

    if (0.0 <= S_L) {
        memcpy(F, FL, sizeof FL);
    } else if (S_L <= 0.0 && 0.0 <= S_star) {
        for (int i = 0; i < 4; i++) F[i] = FL[i] + S_L * (ULs[i] - UL[i]);
    } else if (S_star <= 0.0 && 0.0 <= S_R) {
        for (int i = 0; i < 4; i++) F[i] = FR[i] + S_R * (URs[i] - UR[i]);
    } else if (0.0 >= S_R) {
        memcpy(F, FR, sizeof FR);
    } else {
        for (int i = 0; i < 4; i++) F[i] = 0.0;
    }
}

/* -------------------------------------------------------------- limiters -- */

/* Limiter::denom_tol and xi_right with the parser's fixed parameters
 * omega=0, beta=1, tol=1e-6 (scheme/limiter/limiter/limiter.cpp:26-53,
 * simulation/parser/parser.cpp:231,233). */
static double denom_tol(double denom)
{
    const double tol = 1e-6;
    if (fabs(denom) < tol) {
        if (denom == 0.0) return tol;
        return tol * ((denom > 0) - (denom < 0));
    }
    return denom;
}

static double xi_right(double r)
{
    double num = 2.0 * 1.0;
    double denom = 1.0 - 0.0 + (1.0 + 0.0) * r;
    return num / denom_tol(denom);
}

/* MinBeeLimiter::operator() (scheme/limiter/minbee/minbee.cpp:29-38) and
 * VanLeerLimiter::operator() (scheme/limiter/vanleer/vanleer.cpp:29-36). */
double hfo_limiter(int kind, double r)
{
    if (kind == HFO_MINBEE) {
        if (r <= 0.0) return 0.0;
        if ((0.0 < r) && (r <= 1.0)) return r;
        return smin(1.0, xi_right(r));
    }
    if (r <= 0.0) return 0.0;
    return smin(((2.0 * r) / (1.0 + r)), xi_right(r));
}

/* The tolerance clamp of MUSCLHancockScheme::calculate_slope_limiters_x/_y,
 * scheme/muscl_hancock/muscl_hancock.cpp:8-11,78-94. */
static inline double clamp_delta(double d)
{
    const double tol = 1e-6;
    if (fabs(d) <= tol) {
        if (d == 0.0) return tol;
        return tol * (double)((d > 0) - (d < 0));
    }
    return d;
}

/* ------------------------------------------------------------------ mesh -- */

hfo_mesh* hfo_create(int nblocks, const int* idesc, const double* ddesc,
                     int scheme, int limiter, double gamma, double cfl)
{
    hfo_mesh* m = (hfo_mesh*)calloc(1, sizeof *m);
    m->nb = nblocks;
    m->g = scheme == HFO_MUSCL ? 2 : 1;     /* parser.cpp:275-288 */
    m->scheme = scheme;
    m->limiter = limiter;
    m->gamma = gamma;
    m->cfl = cfl;
    m->steps = 1;
    m->b = (blk_t*)calloc((size_t)nblocks, sizeof(blk_t));
    for (int b = 0; b < nblocks; b++) {
        blk_t* B = &m->b[b];
        B->nx = idesc[10 * b + 0];
        B->ny = idesc[10 * b + 1];
        for (int s = 0; s < 4; s++) {
            B->kind[s] = idesc[10 * b + 2 + s];
            B->nbr[s] = idesc[10 * b + 6 + s];
            for (int k = 0; k < 4; k++) B->state[s][k] = ddesc[20 * b + 4 + 4 * s + k];
        }
        B->xl = ddesc[20 * b + 0];
        B->xr = ddesc[20 * b + 1];
        B->yb = ddesc[20 * b + 2];
        B->yt = ddesc[20 * b + 3];
        B->tx = B->nx + 2 * m->g;
        B->ty = B->ny + 2 * m->g;
        /* MeshBlock::calculate_dx/dy, mesh/mesh_block/mesh_block.cpp:227-236 */
        B->dx = (B->xr - B->xl) / B->nx;
        B->dy = (B->yt - B->yb) / B->ny;
        B->U = (double*)calloc((size_t)4 * B->tx * B->ty, sizeof(double));
    }
    return m;
}

void hfo_destroy(hfo_mesh* m)
{
    if (!m) return;
    for (int b = 0; b < m->nb; b++) free(m->b[b].U);
    free(m->b);
    free(m->log);
    free(m);
}

int hfo_ghost(const hfo_mesh* m) { return m->g; }
unsigned hfo_calls(const hfo_mesh* m) { return m->steps - 1; }

void hfo_set_fields(hfo_mesh* m, int b, const double* U)
{
    blk_t* B = &m->b[b];
    memcpy(B->U, U, sizeof(double) * 4 * B->tx * B->ty);
}

void hfo_get_fields(const hfo_mesh* m, int b, double* U)
{
    const blk_t* B = &m->b[b];
    memcpy(U, B->U, sizeof(double) * 4 * B->tx * B->ty);
}

static inline void copy_cell(blk_t* D, int di, int dj, const blk_t* S, int si, int sj)
{
    for (int v = 0; v < 4; v++) AT(D, v, di, dj) = AT(S, v, si, sj);
}

/* One side of one block.  The loops follow the four BC classes:
 * Transmissive scheme/boundary/transmissive/transmissive_boundary.cpp:25-104,
 * Reflective   scheme/boundary/reflective/reflective_boundary.cpp:25-120,
 * Dirichlet    scheme/boundary/dirichlet/dirichlet_boundary.cpp:25-104,
 * Internal     scheme/boundary/internal/internal_boundary.cpp:27-108.
 * Every side covers the full padded extent, corners included. */
static void apply_side(hfo_mesh* m, int b, int side)
{
    blk_t* B = &m->b[b];
    const int g = m->g, tx = B->tx, ty = B->ty, kind = B->kind[side];
    const blk_t* E = kind == HFO_INTERNAL ? &m->b[B->nbr[side]] : NULL;
    const int n = (side == 0 || side == 2) ? tx : ty;
    /* momentum component negated by a Reflective side */
    const int nv = (side == 0 || side == 2) ? 2 : 1;

    for (int k = 0; k < n; k++) {
        /* (dst, src) pairs in the order the reference writes them */
        int di[2], dj[2], si[2], sj[2], cnt = g;
        for (int l = 0; l < g; l++) {
            /* l = 0 is the ghost layer next to the interior */
            int gi, gj;      /* ghost cell */
            int mi, mj;      /* mirror cell inside the own block */
            int ei = 0, ej = 0; /* source in the neighbour (Internal) */
            switch (side) {
            case 0: /* bottom */
                gi = k; gj = g - 1 - l; mi = k; mj = g + l;
                if (E) { ei = k; ej = E->ty - g - 1 - l; }
                break;
            case 1: /* right */
                gi = tx - g + l; gj = k; mi = tx - g - 1 - l; mj = k;
                if (E) { ei = g + l; ej = k; }
                break;
            case 2: /* top */
                gi = k; gj = ty - g + l; mi = k; mj = ty - g - 1 - l;
                if (E) { ei = k; ej = g + l; }
                break;
            default: /* left */
                gi = g - 1 - l; gj = k; mi = g + l; mj = k;
                if (E) { ei = E->tx - g - 1 - l; ej = k; }
                break;
            }
            di[l] = gi; dj[l] = gj;
            if (E) { si[l] = ei; sj[l] = ej; }
            else { si[l] = mi; sj[l] = mj; }
        }
        /* Internal right/top write the outer layer first (internal_boundary.cpp:62-63,104-105) */
        int order[2] = {0, 1};
        if (kind == HFO_INTERNAL && g == 2 && (side == 1 || side == 2)) { order[0] = 1; order[1] = 0; }
        for (int q = 0; q < cnt; q++) {
            int l = order[q];
            if (kind == HFO_DIRICHLET) {
                for (int v = 0; v < 4; v++) AT(B, v, di[l], dj[l]) = B->state[side][v];
            } else if (kind == HFO_INTERNAL) {
                copy_cell(B, di[l], dj[l], E, si[l], sj[l]);
            } else {
                copy_cell(B, di[l], dj[l], B, si[l], sj[l]);
            }
        }
        if (kind == HFO_REFLECTIVE)
            for (int q = 0; q < cnt; q++) AT(B, nv, di[q], dj[q]) *= -1.0;
    }
}

/* Mesh::apply_bcs, mesh/mesh/mesh.cpp:31-39: per block, bottom, right, top, left. */
void hfo_apply_bcs(hfo_mesh* m)
{
    for (int b = 0; b < m->nb; b++)
        for (int s = 0; s < 4; s++) apply_side(m, b, s);
}

/* -------------------------------------------------------------- timestep -- */

static void log_dt(hfo_mesh* m, double dt)
{
    if (m->nlog == m->caplog) {
        m->caplog = m->caplog ? 2 * m->caplog : 1024;
        m->log = (double*)realloc(m->log, sizeof(double) * m->caplog);
    }
    m->log[m->nlog++] = dt;
}

/* TimeStep::operator(), scheme/timestep/timestep.cpp:29-68: minimum over all
 * blocks and all padded cells (ghosts and corners included); the first ten
 * calls use 0.2*cfl (:60-65,72,75).  Minimum is exact, so the per-thread
 * partials combine to the serial result. */
double hfo_time_step(hfo_mesh* m)
{
    double dtx = 1e16, dty = 1e16;
    for (int b = 0; b < m->nb; b++) {
        const blk_t* B = &m->b[b];
        const double dx = B->dx, dy = B->dy, gamma = m->gamma;
        #pragma omp parallel num_threads(g_threads)
        {
            double lx = 1e16, ly = 1e16;
            #pragma omp for schedule(static)
            for (int j = 0; j < B->ty; j++) {
                for (int i = 0; i < B->tx; i++) {
                    double U[4], W[4];
                    for (int v = 0; v < 4; v++) U[v] = AT(B, v, i, j);
                    hfo_cons_to_prim(gamma, U, W);
                    double c = sqrt(gamma * W[3] / W[0]);
                    lx = smin(lx, dx / (fabs(W[1]) + c));
                    ly = smin(ly, dy / (fabs(W[2]) + c));
                }
            }
            #pragma omp critical
            {
                dtx = smin(dtx, lx);
                dty = smin(dty, ly);
            }
        }
    }
    double final_cfl = (m->steps <= 10) ? 0.2 * m->cfl : m->cfl;
    m->steps++;
    double dt = final_cfl * smin(dtx, dty);
    log_dt(m, dt);
    return dt;
}

int hfo_dt_log(const hfo_mesh* m, double* out, int max)
{
    int n = m->nlog < max ? m->nlog : max;
    if (out) memcpy(out, m->log, sizeof(double) * n);
    return m->nlog;
}

/* --------------------------------------------------------------- schemes -- */

/* GodunovScheme::operator(), scheme/godunov/godunov.cpp:27-118.  Fluxes on
 * ghost rows/columns (:53-87) only feed discarded updates, so the oracle
 * evaluates the faces of interior rows/columns only. */
static void godunov_residual(const hfo_mesh* m, const blk_t* B, double* L)
{
    const int tx = B->tx, ty = B->ty, nx = B->nx, ny = B->ny;
    const double gamma = m->gamma;
    double* Fx = (double*)malloc(sizeof(double) * 4 * (size_t)(nx + 1) * ty);
    double* Fy = (double*)malloc(sizeof(double) * 4 * (size_t)tx * (ny + 1));
    memset(L, 0, sizeof(double) * 4 * (size_t)tx * ty);

    #pragma omp parallel num_threads(g_threads)
    {
        #pragma omp for schedule(static)
        for (int j = 1; j < ty - 1; j++)
            for (int i = 0; i < nx + 1; i++) {
                double UL[4], UR[4], F[4];
                for (int v = 0; v < 4; v++) { UL[v] = AT(B, v, i, j); UR[v] = AT(B, v, i + 1, j); }
                hfo_hllc(gamma, UL, UR, 0, F);
                for (int v = 0; v < 4; v++) Fx[((size_t)j * (nx + 1) + i) * 4 + v] = F[v];
            }
        #pragma omp for schedule(static)
        for (int j = 0; j < ny + 1; j++)
            for (int i = 1; i < tx - 1; i++) {
                double UL[4], UR[4], F[4];
                for (int v = 0; v < 4; v++) { UL[v] = AT(B, v, i, j); UR[v] = AT(B, v, i, j + 1); }
                hfo_hllc(gamma, UL, UR, 1, F);
                for (int v = 0; v < 4; v++) Fy[((size_t)j * tx + i) * 4 + v] = F[v];
            }
    }
    /* calculate_spatial_update, godunov.cpp:93-118 */
    const double one_dx = 1.0 / B->dx, one_dy = 1.0 / B->dy;
    for (int j = 1; j < ty - 1; j++)
        for (int i = 1; i < tx - 1; i++)
            for (int v = 0; v < 4; v++) {
                double fw = Fx[((size_t)j * (nx + 1) + (i - 1)) * 4 + v];
                double fe = Fx[((size_t)j * (nx + 1) + i) * 4 + v];
                double gs = Fy[((size_t)(j - 1) * tx + i) * 4 + v];
                double gn = Fy[((size_t)j * tx + i) * 4 + v];
                L[((size_t)v * ty + j) * tx + i] = (one_dx * (fw - fe)) + (one_dy * (gs - gn));
            }
    free(Fx);
    free(Fy);
}

/* MUSCLHancockScheme::operator(), scheme/muscl_hancock/muscl_hancock.cpp:41-376.
 * rec holds, per padded cell, the four face states (L, R, B, T) x 4 vars. */
static void muscl_residual(hfo_mesh* m, int b, double* L)
{
    blk_t* B = &m->b[b];
    const int tx = B->tx, ty = B->ty, nx = B->nx, ny = B->ny, lim = m->limiter;
    const double gamma = m->gamma;
    double* rec = (double*)calloc((size_t)tx * ty * 16, sizeof(double));
#define REC(i, j, face, v) rec[(((size_t)(j) * tx + (i)) * 4 + (face)) * 4 + (v)]

    /* slopes (:137-206) and boundary extrapolations (:209-240) */
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int j = 1; j < ty - 1; j++)
        for (int i = 1; i < tx - 1; i++)
            for (int v = 0; v < 4; v++) {
                double u0 = AT(B, v, i, j);
                /* x: delta_x_back / delta_x_forward (mesh_block.cpp:167-180) */
                double dl = AT(B, v, i, j) - AT(B, v, i - 1, j);
                double dr = AT(B, v, i + 1, j) - AT(B, v, i, j);
                double r = clamp_delta(dl) / clamp_delta(dr);
                double sx = hfo_limiter(lim, r) * (0.5 * (1.0 + 0.0) * dl + 0.5 * (1.0 - 0.0) * dr);
                /* y: delta_y_back / delta_y_forward (mesh_block.cpp:191-204) */
                double db = AT(B, v, i, j) - AT(B, v, i, j - 1);
                double dt_ = AT(B, v, i, j + 1) - AT(B, v, i, j);
                double ry = clamp_delta(db) / clamp_delta(dt_);
                double sy = hfo_limiter(lim, ry) * (0.5 * (1.0 + 0.0) * db + 0.5 * (1.0 - 0.0) * dt_);
                REC(i, j, 0, v) = u0 - 0.5 * sx;
                REC(i, j, 1, v) = u0 + 0.5 * sx;
                REC(i, j, 2, v) = u0 - 0.5 * sy;
                REC(i, j, 3, v) = u0 + 0.5 * sy;
            }

    /* evolve_boundary_reconstruction (:243-281): the TimeStep is called
     * again here, over the whole mesh in its current state; the predictor is
     * direction-split (x faces use F only, y faces G only). */
    double dt = hfo_time_step(m);
    double half_dt_dx = 0.5 * dt / B->dx;
    double half_dt_dy = 0.5 * dt / B->dy;
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int j = 1; j < ty - 1; j++)
        for (int i = 1; i < tx - 1; i++) {
            double Uf[4][4], Fl[4], Fr[4], Gb[4], Gt[4];
            for (int f = 0; f < 4; f++)
                for (int v = 0; v < 4; v++) Uf[f][v] = REC(i, j, f, v);
            hfo_cons_flux(gamma, Uf[0], 0, Fl);
            hfo_cons_flux(gamma, Uf[1], 0, Fr);
            hfo_cons_flux(gamma, Uf[2], 1, Gb);
            hfo_cons_flux(gamma, Uf[3], 1, Gt);
            for (int v = 0; v < 4; v++) {
                REC(i, j, 0, v) += half_dt_dx * (Fl[v] - Fr[v]);
                REC(i, j, 1, v) += half_dt_dx * (Fl[v] - Fr[v]);
                REC(i, j, 2, v) += half_dt_dy * (Gb[v] - Gt[v]);
                REC(i, j, 3, v) += half_dt_dy * (Gb[v] - Gt[v]);
            }
        }

    /* calculate_fluxes_x (:284-313) and calculate_fluxes_y (:316-345) */
    double* Fx = (double*)malloc(sizeof(double) * 4 * (size_t)(nx + 1) * ny);
    double* Fy = (double*)malloc(sizeof(double) * 4 * (size_t)nx * (ny + 1));
    #pragma omp parallel num_threads(g_threads)
    {
        #pragma omp for schedule(static)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx + 1; i++) {
                double UL[4], UR[4], F[4];
                for (int v = 0; v < 4; v++) { UL[v] = REC(i + 1, j + 2, 1, v); UR[v] = REC(i + 2, j + 2, 0, v); }
                hfo_hllc(gamma, UL, UR, 0, F);
                for (int v = 0; v < 4; v++) Fx[((size_t)j * (nx + 1) + i) * 4 + v] = F[v];
            }
        #pragma omp for schedule(static)
        for (int j = 0; j < ny + 1; j++)
            for (int i = 0; i < nx; i++) {
                double UL[4], UR[4], F[4];
                for (int v = 0; v < 4; v++) { UL[v] = REC(i + 2, j + 1, 3, v); UR[v] = REC(i + 2, j + 2, 2, v); }
                hfo_hllc(gamma, UL, UR, 1, F);
                for (int v = 0; v < 4; v++) Fy[((size_t)j * nx + i) * 4 + v] = F[v];
            }
    }
#undef REC

    /* calculate_spatial_update (:348-376): rdx = 1.0/dx, then a multiply */
    const double one_dx = 1.0 / B->dx, one_dy = 1.0 / B->dy;
    memset(L, 0, sizeof(double) * 4 * (size_t)tx * ty);
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
            for (int v = 0; v < 4; v++) {
                double fw = Fx[((size_t)j * (nx + 1) + i) * 4 + v];
                double fe = Fx[((size_t)j * (nx + 1) + i + 1) * 4 + v];
                double gs = Fy[((size_t)j * nx + i) * 4 + v];
                double gn = Fy[((size_t)(j + 1) * nx + i) * 4 + v];
                L[((size_t)v * ty + (j + 2)) * tx + (i + 2)] = (one_dx * (fw - fe)) + (one_dy * (gs - gn));
            }
    free(Fx);
    free(Fy);
    free(rec);
}

void hfo_scheme_residual(hfo_mesh* m, int b, double* L)
{
    if (m->scheme == HFO_GODUNOV) godunov_residual(m, &m->b[b], L);
    else muscl_residual(m, b, L);
}

/* Simulation::solve step body (simulation/simulation/simulation.cpp:85-110)
 * without the t_end / frame clamps; ForwardEulerODESolver::operator()
 * (scheme/ode/forward_euler/forward_euler.cpp:15-33) updates the whole
 * padded block as (dt*L) + U, blocks in order. */
double hfo_step(hfo_mesh* m)
{
    hfo_apply_bcs(m);
    double dt = hfo_time_step(m);
    for (int b = 0; b < m->nb; b++) {
        blk_t* B = &m->b[b];
        size_t n = (size_t)4 * B->tx * B->ty;
        double* L = (double*)malloc(sizeof(double) * n);
        hfo_scheme_residual(m, b, L);
        for (size_t k = 0; k < n; k++) B->U[k] = dt * L[k] + B->U[k];
        free(L);
    }
    return dt;
}

int hfo_run(hfo_mesh* m, int nsteps, double* dt_steps, double* dt_calls, int max_calls)
{
    int first = m->nlog;
    for (int n = 0; n < nsteps; n++) {
        double dt = hfo_step(m);
        if (dt_steps) dt_steps[n] = dt;
    }
    int made = m->nlog - first;
    if (dt_calls) {
        int k = made < max_calls ? made : max_calls;
        memcpy(dt_calls, m->log + first, sizeof(double) * k);
    }
    return made;
}


/* ===================================================================== */
/* Unstructured quadrilateral-cell Godunov solver (src/fvm_2d_uns).       */
/* ===================================================================== */

struct hfo_uns {
    int n;
    double gamma, cfl, inlet[4];
    int* nbr;          /* [n][4] neighbour cell or boundary code */
    double* len;       /* [n][4] edge length */
    double* cs;        /* [n][4][4] cos, sin of the outward normal angle, then of its negative */
    double* inv_vol;   /* [n] */
    double* dr;        /* [n] incircle radius (shortest edge) */
    double* cen;       /* [n][2] */
    double* vol;       /* [n] */
    double* U;         /* [n][4] */
    double* L;         /* [n][4] scratch */
    unsigned steps;
};

/* Edge(start, end) members (geometry/edge/edge.cpp:12-238, tol 1e-12). */
static void uns_edge(const double* a, const double* b, double* len, double* cs)
{
    const double px = b[0] - a[0], py = b[1] - a[1];           /* calculate_point_vector */
    *len = sqrt((0.0 + px * px) + py * py);                     /* l2_norm = sqrt(dot) */
    double n0 = py, n1 = -1.0 * px;                             /* calculate_outward_unit_normal */
    const double nl = sqrt(n0 * n0 + n1 * n1);
    n0 /= nl;
    n1 /= nl;
    const double deg = atan2(n1, n0) * 180.0 / M_PI;            /* calculate_outward_normal_angle */
    const double r = deg * (M_PI / 180.0);                      /* rotate_flow_values (euler_ideal_gas.cpp:155) */
    const double rb = (-1.0 * deg) * (M_PI / 180.0);            /* the rotation back (godunov.cpp:62) */
    /* rotate_flow_values takes cos and sin of one angle; the reference's
     * build (g++ -O3, build/fvm_2d_uns/makefile) merges the pair into one
     * glibc sincos call, whose results differ from separate cos and sin in
     * the last bit for some angles.  Called explicitly here. */
    sincos(r, &cs[1], &cs[0]);
    sincos(rb, &cs[3], &cs[2]);
}

static int uns_opposite(const double* a0, const double* a1, const double* b0, const double* b1)
{
    /* Edge::opposite_edges_equal (edge.cpp:47-66): rhs start == end, rhs end == start */
    const double tol = 1e-12;
    return fabs(b0[0] - a1[0]) < tol && fabs(b0[1] - a1[1]) < tol && fabs(b1[0] - a0[0]) < tol &&
           fabs(b1[1] - a0[1]) < tol;
}

hfo_uns* hfo_uns_create(int n, const double* V, const int* codes, double gamma, double cfl, const double* inlet)
{
    hfo_uns* m = (hfo_uns*)calloc(1, sizeof(hfo_uns));
    m->n = n;
    m->gamma = gamma;
    m->cfl = cfl;
    for (int k = 0; k < 4; k++) m->inlet[k] = inlet[k];
    m->nbr = (int*)malloc(sizeof(int) * 4 * n);
    m->len = (double*)malloc(sizeof(double) * 4 * n);
    m->cs = (double*)malloc(sizeof(double) * 16 * n);
    m->inv_vol = (double*)malloc(sizeof(double) * n);
    m->dr = (double*)malloc(sizeof(double) * n);
    m->cen = (double*)malloc(sizeof(double) * 2 * n);
    m->vol = (double*)malloc(sizeof(double) * n);
    m->U = (double*)calloc(4 * (size_t)n, sizeof(double));
    m->L = (double*)calloc(4 * (size_t)n, sizeof(double));
    m->steps = 1;
    for (int c = 0; c < n; c++) {
        const double* q = V + 8 * c;
        /* Polygon::calculate_volume / calculate_centroid (polygon.cpp:162-210) */
        double vol = 0.0;
        for (int i = 0; i < 4; i++) {
            const int j = (i + 1) % 4;
            vol += (q[2 * i] * q[2 * j + 1] - q[2 * j] * q[2 * i + 1]);
        }
        vol *= 0.5;
        double cx = 0.0, cy = 0.0;
        for (int i = 0; i < 4; i++) {
            const int j = (i + 1) % 4;
            const double second = (q[2 * i] * q[2 * j + 1] - q[2 * j] * q[2 * i + 1]);
            cx += ((q[2 * i] + q[2 * j]) * second);
            cy += ((q[2 * i + 1] + q[2 * j + 1]) * second);
        }
        m->vol[c] = vol;
        m->cen[2 * c] = cx / (6.0 * vol);
        m->cen[2 * c + 1] = cy / (6.0 * vol);
        m->inv_vol[c] = 1.0 / vol;                               /* godunov.cpp:71 */
        double shortest = 1e16;                                  /* Polygon::radius_incircle, 4 sides */
        for (int e = 0; e < 4; e++) {
            uns_edge(q + 2 * e, q + 2 * ((e + 1) % 4), &m->len[4 * c + e], &m->cs[16 * c + 4 * e]);
            shortest = (shortest < m->len[4 * c + e]) ? shortest : m->len[4 * c + e];   /* std::min(len, shortest) */
        }
        m->dr[c] = shortest;
    }
    /* LoadStructuredMesh::load_structured_mesh (mesh_structured.cpp:93-115):
     * the edge's boundary code, replaced by the last cell (in cell, edge
     * order) that has the same edge in the opposite direction */
    for (int c = 0; c < n; c++)
        for (int e = 0; e < 4; e++) {
            const double* a0 = V + 8 * c + 2 * e;
            const double* a1 = V + 8 * c + 2 * ((e + 1) % 4);
            int nb = codes[4 * c + e];
            for (int d = 0; d < n; d++)
                for (int f = 0; f < 4; f++)
                    if (uns_opposite(a0, a1, V + 8 * d + 2 * f, V + 8 * d + 2 * ((f + 1) % 4))) nb = d;
            m->nbr[4 * c + e] = nb;
        }
    return m;
}

void hfo_uns_destroy(hfo_uns* m)
{
    if (!m) return;
    free(m->nbr);
    free(m->len);
    free(m->cs);
    free(m->inv_vol);
    free(m->dr);
    free(m->cen);
    free(m->vol);
    free(m->U);
    free(m->L);
    free(m);
}

void hfo_uns_connectivity(const hfo_uns* m, int* out) { memcpy(out, m->nbr, sizeof(int) * 4 * m->n); }

void hfo_uns_geometry(const hfo_uns* m, double* centroids, double* volumes)
{
    memcpy(centroids, m->cen, sizeof(double) * 2 * m->n);
    memcpy(volumes, m->vol, sizeof(double) * m->n);
}

void hfo_uns_set_fields(hfo_uns* m, const double* U) { memcpy(m->U, U, sizeof(double) * 4 * m->n); }
void hfo_uns_get_fields(const hfo_uns* m, double* U) { memcpy(U, m->U, sizeof(double) * 4 * m->n); }

/* TimeStep::operator() (scheme/timestep/timestep.cpp:29-62) */
double hfo_uns_time_step(hfo_uns* m)
{
    double dtr = 1e16;
    for (int c = 0; c < m->n; c++) {
        double W[4];
        hfo_cons_to_prim(m->gamma, m->U + 4 * c, W);
        const double u = W[1], v = W[2];
        const double a = sqrt(m->gamma * W[3] / W[0]);
        dtr = smin(dtr, m->dr[c] / (fabs(sqrt(u * u + v * v)) + a));
    }
    const double cfl = (m->steps <= 10) ? 0.2 * m->cfl : m->cfl;
    m->steps++;
    return cfl * dtr;
}

/* EulerIdealGasModel::rotate_flow_values (euler_ideal_gas.cpp:152-164) */
static void uns_rotate(double c, double s, const double* W, double* R)
{
    R[0] = W[0];
    R[1] = c * W[1] + s * W[2];
    R[2] = -1.0 * s * W[1] + c * W[2];
    R[3] = W[3];
}

/* GodunovScheme::operator() (scheme/godunov/godunov.cpp:31-77) with the
 * boundary states of BoundaryConditions (scheme/boundary/boundary.cpp:29-69)
 * and ForwardEulerODESolver (scheme/ode/forward_euler/forward_euler.cpp:16-30). */
void hfo_uns_advance(hfo_uns* m, double dt)
{
    const int n = m->n;
    for (int c = 0; c < n; c++) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        const double* U = m->U + 4 * c;
        for (int e = 0; e < 4; e++) {
            const int nb = m->nbr[4 * c + e];
            const double* cs = m->cs + 16 * c + 4 * e;
            double V[4], rU[4], rV[4], F[4], f[4];
            if (nb >= 0) {
                memcpy(V, m->U + 4 * nb, sizeof V);
            } else if (nb == -1) {                  /* reflective */
                double t[4];
                uns_rotate(cs[0], cs[1], U, t);
                t[1] *= -1.0;
                uns_rotate(cs[2], cs[3], t, V);
            } else if (nb == -2) {                  /* transmissive */
                memcpy(V, U, sizeof V);
            } else {                                /* inlet */
                memcpy(V, m->inlet, sizeof V);
            }
            uns_rotate(cs[0], cs[1], U, rU);
            uns_rotate(cs[0], cs[1], V, rV);
            hfo_hllc(m->gamma, rU, rV, 0, F);
            uns_rotate(cs[2], cs[3], F, f);
            for (int k = 0; k < 4; k++) {
                f[k] *= m->len[4 * c + e];
                acc[k] += f[k];
            }
        }
        const double iv = m->inv_vol[c];
        for (int k = 0; k < 4; k++) {
            acc[k] *= iv;
            acc[k] *= -1.0;
            m->L[4 * c + k] = acc[k];
        }
    }
    for (int i = 0; i < 4 * n; i++) {
        double x = m->L[i];
        x *= dt;
        x += m->U[i];
        m->U[i] = x;
    }
}
