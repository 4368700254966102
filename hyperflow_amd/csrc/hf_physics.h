// hf_physics.h -- per-cell and per-face arithmetic of the Euler step for gfx950.
//
// Every expression keeps the reference's evaluation order (binary64 with the
// FP64 policy, compiled with -ffp-contract=off), so the FP64 kernels are
// bit-identical to the reference CPU path.  Paths below are relative to the
// reference checkout under src/fvm_2d_cart/.
//
// Deliberate, bit-preserving differences from the reference's code shape:
//  * primitives, E and the sound speed of a state are computed once and
//    shared by its flux and the HLLC star state (same formula, same operands);
//  * HLLC evaluates only the star state of the selected branch;
//  * the limiters evaluate one division (a correctly rounded quotient is
//    monotone in the numerator, so min(2r/(1+r), 2/(1+r)) = min(2r,2)/(1+r)).
#pragma once
#include <hip/hip_runtime.h>

namespace hf {

// Exact binary64: IEEE division and the correctly rounded square root.
struct MathF64 {
    typedef double R;
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    static __device__ __forceinline__ double sqrt(double x) { return __builtin_sqrt(x); }
    static __device__ __forceinline__ double fabs(double x) { return __builtin_fabs(x); }
    static constexpr double tol = 1e-6;
    static constexpr bool fast = false;   // the reference's expressions, in its order
};

// Bulk FP32: hardware reciprocal and square root (about 1 ulp each).
struct MathF32 {
    typedef float R;
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float fabs(float x) { return __builtin_fabsf(x); }
    static constexpr float tol = 1e-6f;
    static constexpr bool fast = true;    // the FP32 formulation of namespace f32 below
};

// std::min / std::max (NaN handling of the reference:
// scheme/riemann/riemann_hllc/riemann_hllc.cpp:146-147, timestep.cpp:51-52).
template <class R> __device__ __forceinline__ R smin(R a, R b) { return (b < a) ? b : a; }
template <class R> __device__ __forceinline__ R smax(R a, R b) { return (a < b) ? b : a; }

// A conserved state with its primitives, total energy (re-derived from the
// primitives) and sound speed.
template <class R>
struct State {
    R U0, U1, U2, U3;   // rho, rho*u, rho*v, E
    R u, v, p, E, a;
};

// EulerIdealGasModel::cons_to_prim (model/euler_ideal_gas/euler_ideal_gas.cpp:76-92)
// and prim_total_energy / prim_internal_energy (:32-46).
template <class M>
__device__ __forceinline__ void prims(typename M::R U0, typename M::R U1, typename M::R U2, typename M::R U3,
                                      typename M::R g1, typename M::R& u, typename M::R& v,
                                      typename M::R& p, typename M::R& E)
{
    typedef typename M::R R;
    u = M::div(U1, U0);
    v = M::div(U2, U0);
    const R q2 = u * u + v * v;
    p = g1 * (U3 - R(0.5) * U0 * q2);
    const R e = M::div(p, U0 * g1);
    E = U0 * (R(0.5) * q2 + e);
}

// prim_speed_of_sound (euler_ideal_gas.cpp:60-64), also used by
// calc_wave_speeds_direct_min_max (riemann_hllc.cpp:130-131).
template <class M>
__device__ __forceinline__ typename M::R sound(typename M::R gamma, typename M::R p, typename M::R rho)
{
    return M::sqrt(M::div(gamma * p, rho));
}

// ------------------------------------------------------------ FP32 bulk --
// The FP32 path owes no bit-for-bit contract, so it evaluates the same
// scheme in fewer operations (results agree with the FP64 path to FP32
// rounding):
//  * one reciprocal of rho per state, shared by u, v and the sound speed;
//  * the state's own total energy E where the reference re-derives it from
//    the primitives (rho (q^2/2 + p/(rho (gamma-1))) equals E up to rounding);
//  * rho u and rho v are the conserved momenta themselves in the fluxes;
//  * the limiters with one reciprocal (the ratio and xi(r) folded into one
//    quotient), HLLC's star state with two, the flux of the upwind side only;
//  * multiply-adds written as explicit fmaf.  The build disables contraction,
//    so no other operation is fused and every copy of this code (warm-up and
//    loop body of a sweep) rounds the same way: results do not depend on how
//    the rows are split into work items.
namespace f32 {

__device__ __forceinline__ float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// the limiter ratio's tolerance clamp (muscl_hancock.cpp:8-11,78-94):
// |d| <= 1e-6 becomes +-1e-6 (zero becomes +1e-6)
__device__ __forceinline__ float clamp_tol(float d)
{
    const float tol = 1e-6f;
    return (__builtin_fabsf(d) <= tol) ? ((d >= 0.0f) ? tol : -tol) : d;
}

// u, v, p of a conserved state; ri = 1/rho
__device__ __forceinline__ void prims(float U0, float U1, float U2, float U3, float g1, float& ri, float& u, float& v,
                                      float& p)
{
    ri = rcp(U0);
    u = U1 * ri;
    v = U2 * ri;
    p = g1 * fma(-0.5f, fma(U1, u, U2 * v), U3);
}

__device__ __forceinline__ State<float> make_state(const float U[4], float gamma, float g1)
{
    State<float> s;
    s.U0 = U[0]; s.U1 = U[1]; s.U2 = U[2]; s.U3 = U[3];
    float ri;
    prims(U[0], U[1], U[2], U[3], g1, ri, s.u, s.v, s.p);
    s.E = U[3];
    s.a = __builtin_amdgcn_sqrtf(gamma * s.p * ri);
    return s;
}

// Euler flux normal to the face: X (x-face) or y-face, E = the state's own energy
template <bool X>
__device__ __forceinline__ void flux(float U0, float U1, float U2, float U3, float u, float v, float p, float F[4])
{
    (void)U0;
    if (X) {
        F[0] = U1;
        F[1] = fma(U1, u, p);
        F[2] = U1 * v;
        F[3] = u * (U3 + p);
    } else {
        F[0] = U2;
        F[1] = U2 * u;
        F[2] = fma(U2, v, p);
        F[3] = v * (U3 + p);
    }
}

// HLLC with the Davis wave speeds, Toro's star state rewritten with
// m_K = rho_K (S_K - q_K):  S* = (p_R - p_L + m_L q_L - m_R q_R) / (m_L - m_R),
// U*_K = [m_K, m_K S*, m_K t_K, (S_K - q_K) E_K + (S* - q_K)(m_K S* + p_K)] / (S_K - S*).
// The branch order of the reference: 0 <= S_L full left flux, then the left
// and right star fluxes, 0 >= S_R full right flux, otherwise (NaN) zero.
template <bool X>
__device__ __forceinline__ void hllc(const State<float>& L, const State<float>& Rs, float F[4])
{
    const float qL = X ? L.u : L.v, qR = X ? Rs.u : Rs.v;
    const float S_L = smin(qL - L.a, qR - Rs.a);
    const float S_R = smax(qL + L.a, qR + Rs.a);
    const float mL = L.U0 * (S_L - qL), mR = Rs.U0 * (S_R - qR);
    const float Ss = fma(mL, qL, fma(-mR, qR, Rs.p - L.p)) * rcp(mL - mR);

    // Branch choice of the reference for finite speeds: S_L >= 0 takes the
    // full left flux; otherwise S* >= 0 the left star flux; otherwise the
    // right side, its star flux if S_R >= 0, else the full right flux.  So
    // K = R exactly when S_L < 0 and S* < 0, and the star term's weight
    // S_K (zero for a full flux) is min(S_L, 0) on the left, max(S_R, 0) on
    // the right.  A NaN speed gives the zero flux, as in the reference.
    const bool useR = (S_L < 0.0f) & (Ss < 0.0f);
    const float wL = (S_L < 0.0f) ? S_L : 0.0f, wR = (S_R > 0.0f) ? S_R : 0.0f;
    const float chk = S_L + S_R + Ss;
    const bool none = !(chk == chk);
    const float dK = useR ? Rs.U0 : L.U0;
    const float U1K = useR ? Rs.U1 : L.U1;
    const float U2K = useR ? Rs.U2 : L.U2;
    const float EK = useR ? Rs.U3 : L.U3;
    const float uK = useR ? Rs.u : L.u;
    const float vK = useR ? Rs.v : L.v;
    const float pK = useR ? Rs.p : L.p;
    const float SK = useR ? S_R : S_L;
    const float mK = useR ? mR : mL;
    const float qK = X ? uK : vK, tK = X ? vK : uK;
    const float UnK = X ? U1K : U2K, UtK = X ? U2K : U1K;

    float FK[4];
    flux<X>(dK, U1K, U2K, EK, uK, vK, pK, FK);
    const float rS = rcp(SK - Ss);
    const float us0 = mK * rS;
    const float us3 = rS * fma(SK - qK, EK, (Ss - qK) * fma(mK, Ss, pK));
    const float w = useR ? wR : wL;
    const bool star = w != 0.0f;   // a full flux never sees the star state (it may not be finite)
    const float d0 = us0 - dK;
    const float dn = fma(us0, Ss, -UnK);
    const float dt = fma(us0, tK, -UtK);
    const float d3 = us3 - EK;
    const float F0 = star ? fma(w, d0, FK[0]) : FK[0];
    const float Fn = star ? fma(w, X ? dn : dt, FK[1]) : FK[1];
    const float Ft = star ? fma(w, X ? dt : dn, FK[2]) : FK[2];
    const float F3 = star ? fma(w, d3, FK[3]) : FK[3];
    F[0] = none ? 0.0f : F0;
    F[1] = none ? 0.0f : Fn;
    F[2] = none ? 0.0f : Ft;
    F[3] = none ? 0.0f : F3;
}

// Limited slope xi(r) * (dl + dr) / 2 with r from the clamped deltas
// (muscl_hancock.cpp:67-206), one reciprocal.  For r > 0 (same signs):
//  MinBee  r <= 1: xi = r              -> (dl+dr)/2 * dlc / drc
//          r >  1: xi = 2/(1+r)        -> (dl+dr) * drc / (drc + dlc)
//  VanLeer xi = 2 min(r, 1)/(1+r)      -> (dl+dr) * m / (dlc + drc), m = dlc or drc, whichever is
//          smaller in magnitude (both have the sign of drc)
template <int LIM>
__device__ __forceinline__ float slope(float um, float uc, float up)
{
    const float dl = uc - um, dr = up - uc;
    const float dlc = clamp_tol(dl), drc = clamp_tol(dr);
    // r > 0: the clamped deltas have one sign (they are never zero, and their
    // product stays far above the FP32 underflow)
    const bool same = dlc * drc > 0.0f;
    const bool le1 = __builtin_fabsf(dlc) <= __builtin_fabsf(drc);   // |r| <= 1
    float num, den;
    if (LIM == 0) {
        num = le1 ? 0.5f * dlc : drc;
        den = le1 ? drc : drc + dlc;
    } else {
        num = le1 ? dlc : drc;
        den = dlc + drc;
    }
    const float s = (dl + dr) * num * rcp(den);
    return same ? s : 0.0f;
}

// One direction of MUSCL-Hancock reconstruction of a cell: face states
// U -+ s/2 and the split half-step predictor face += h (F(lo) - F(hi)).
template <int LIM, bool X>
__device__ __forceinline__ void reconstruct(const float um[4], const float uc[4], const float up[4], float h, float g1,
                                            float lo[4], float hi[4])
{
    for (int k = 0; k < 4; k++) {
        const float s = slope<LIM>(um[k], uc[k], up[k]);
        lo[k] = fma(-0.5f, s, uc[k]);
        hi[k] = fma(0.5f, s, uc[k]);
    }
    float ri, ul, vl, pl, ur, vr, pr;
    prims(lo[0], lo[1], lo[2], lo[3], g1, ri, ul, vl, pl);
    prims(hi[0], hi[1], hi[2], hi[3], g1, ri, ur, vr, pr);
    float Fl[4], Fr[4];
    flux<X>(lo[0], lo[1], lo[2], lo[3], ul, vl, pl, Fl);
    flux<X>(hi[0], hi[1], hi[2], hi[3], ur, vr, pr, Fr);
    for (int k = 0; k < 4; k++) {
        const float d = Fl[k] - Fr[k];
        lo[k] = fma(h, d, lo[k]);
        hi[k] = fma(h, d, hi[k]);
    }
}

// |u| + c and |v| + c of a state (TimeStep's per-cell terms)
__device__ __forceinline__ void wave_speeds(float U0, float U1, float U2, float U3, float gamma, float g1, double& Dx,
                                            double& Dy)
{
    float ri, u, v, p;
    prims(U0, U1, U2, U3, g1, ri, u, v, p);
    const float c = __builtin_amdgcn_sqrtf(gamma * p * ri);
    Dx = (double)(__builtin_fabsf(u) + c);
    Dy = (double)(__builtin_fabsf(v) + c);
}

}  // namespace f32

template <class M>
__device__ __forceinline__ State<typename M::R> make_state(const typename M::R U[4], typename M::R gamma,
                                                           typename M::R g1)
{
    if constexpr (M::fast) {
        return f32::make_state(U, gamma, g1);
    } else {
        State<typename M::R> s;
        s.U0 = U[0]; s.U1 = U[1]; s.U2 = U[2]; s.U3 = U[3];
        prims<M>(U[0], U[1], U[2], U[3], g1, s.u, s.v, s.p, s.E);
        s.a = sound<M>(gamma, s.p, U[0]);
        return s;
    }
}

// cons_flux_x / cons_flux_y (euler_ideal_gas.cpp:113-149) from precomputed
// primitives: [rho u, rho u u + p, rho u v, u (E + p)] and the y mirror.
template <class R, bool X>
__device__ __forceinline__ void flux(R rho, R u, R v, R p, R E, R F[4])
{
    if (X) {
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

// HLLCEulerRiemannSolver::calc_hllc_flux (riemann_hllc.cpp:169-212) with the
// Davis wave speeds (:127-150), S* (:67-97), U*_K (:37-63), F*_K (:153-164).
// NaN or inconsistent states fall through to a zero flux (:207-210).
template <class M, bool X>
__device__ __forceinline__ void hllc(const State<typename M::R>& L, const State<typename M::R>& Rs,
                                     typename M::R F[4])
{
    typedef typename M::R R;
    if constexpr (M::fast) {
        f32::hllc<X>(L, Rs, F);
        return;
    }
    const R qL = X ? L.u : L.v, qR = X ? Rs.u : Rs.v;
    const R S_L = smin(qL - L.a, qR - Rs.a);
    const R S_R = smax(qL + L.a, qR + Rs.a);
    const R num = Rs.p - L.p + L.U0 * qL * (S_L - qL) - Rs.U0 * qR * (S_R - qR);
    const R den = L.U0 * (S_L - qL) - Rs.U0 * (S_R - qR);
    const R Ss = M::div(num, den);

    const bool c1 = R(0) <= S_L;
    const bool c2 = (S_L <= R(0)) && (R(0) <= Ss);
    const bool c3 = (Ss <= R(0)) && (R(0) <= S_R);
    const bool c4 = R(0) >= S_R;
    const bool useR = !c1 && !c2 && c3;   // F*_R branch; otherwise the star state of L

    // the side whose star state is needed
    const R dK = useR ? Rs.U0 : L.U0;
    const R uK = useR ? Rs.u : L.u;
    const R vK = useR ? Rs.v : L.v;
    const R pK = useR ? Rs.p : L.p;
    const R EK = useR ? Rs.E : L.E;
    const R SK = useR ? S_R : S_L;
    const R qK = X ? uK : vK;
    const R tK = X ? vK : uK;
    const R coef = dK * M::div(SK - qK, SK - Ss);
    const R us0 = coef;
    const R usn = coef * Ss;
    const R ust = coef * tK;
    const R us3 = coef * (M::div(EK, dK) + (Ss - qK) * (Ss + M::div(pK, dK * (SK - qK))));

    R FL[4], FR[4];
    flux<R, X>(L.U0, L.u, L.v, L.p, L.E, FL);
    flux<R, X>(Rs.U0, Rs.u, Rs.v, Rs.p, Rs.E, FR);
    const R K0 = useR ? Rs.U0 : L.U0;
    const R K1 = useR ? Rs.U1 : L.U1;
    const R K2 = useR ? Rs.U2 : L.U2;
    const R K3 = useR ? Rs.U3 : L.U3;
    R FK[4];
    for (int i = 0; i < 4; i++) FK[i] = useR ? FR[i] : FL[i];
    R Fs[4];
    Fs[0] = FK[0] + SK * (us0 - K0);
    Fs[1] = FK[1] + SK * ((X ? usn : ust) - K1);
    Fs[2] = FK[2] + SK * ((X ? ust : usn) - K2);
    Fs[3] = FK[3] + SK * (us3 - K3);

    for (int i = 0; i < 4; i++) {
        R f = R(0);
        if (c1) f = FL[i];
        else if (c2 || c3) f = Fs[i];
        else if (c4) f = FR[i];
        F[i] = f;
    }
}

// Tolerance clamp of the limiter ratio (scheme/muscl_hancock/muscl_hancock.cpp:8-11,78-94).
template <class M>
__device__ __forceinline__ typename M::R clamp_delta(typename M::R d)
{
    typedef typename M::R R;
    const R tol = R(M::tol);
    if (M::fabs(d) <= tol) return (d == R(0)) ? tol : ((d > R(0)) ? tol : -tol);
    return d;
}

// MinBeeLimiter::operator() (scheme/limiter/minbee/minbee.cpp:29-38) and
// VanLeerLimiter::operator() (scheme/limiter/vanleer/vanleer.cpp:29-36) with
// omega=0, beta=1 (parser.cpp:231,233); xi_right = 2/(1+r) (limiter.cpp:48-53),
// whose denominator guard never triggers on the evaluated branches.
template <class M, int LIM>
__device__ __forceinline__ typename M::R limiter(typename M::R r)
{
    typedef typename M::R R;
    if (LIM == 0) {
        const R x = M::div(R(2), R(1) + r);
        return (r <= R(0)) ? R(0) : ((r <= R(1)) ? r : ((x < R(1)) ? x : R(1)));
    } else {
        // min(2r/(1+r), 2/(1+r)) with one division; r = +inf gives inf/inf = NaN
        // exactly as std::min(NaN, 0) does in the reference.
        const R inf = __builtin_huge_val();
        const R num = (r < R(1)) ? R(2) * r : ((r == inf) ? inf : R(2));
        const R x = M::div(num, R(1) + r);
        return (r <= R(0)) ? R(0) : x;
    }
}

// Limited slope of one conserved variable (muscl_hancock.cpp:137-206):
// xi(r) * (0.5*(1+omega)*dl + 0.5*(1-omega)*dr) with unclamped deltas.
template <class M, int LIM>
__device__ __forceinline__ typename M::R slope(typename M::R um, typename M::R uc, typename M::R up)
{
    typedef typename M::R R;
    const R dl = uc - um;
    const R dr = up - uc;
    const R r = M::div(clamp_delta<M>(dl), clamp_delta<M>(dr));
    return limiter<M, LIM>(r) * (R(0.5) * dl + R(0.5) * dr);
}

// One direction of MUSCL-Hancock reconstruction for one cell: boundary
// extrapolation (muscl_hancock.cpp:209-240) and the direction-split half-step
// predictor (:243-281): lo/hi are the face states (L,R) or (B,T).
template <class M, int LIM, bool X>
__device__ __forceinline__ void reconstruct(const typename M::R um[4], const typename M::R uc[4],
                                            const typename M::R up[4], typename M::R h, typename M::R g1,
                                            typename M::R lo[4], typename M::R hi[4])
{
    typedef typename M::R R;
    if constexpr (M::fast) {
        f32::reconstruct<LIM, X>(um, uc, up, h, g1, lo, hi);
        return;
    }
    for (int k = 0; k < 4; k++) {
        const R s = slope<M, LIM>(um[k], uc[k], up[k]);
        lo[k] = uc[k] - R(0.5) * s;
        hi[k] = uc[k] + R(0.5) * s;
    }
    R ul, vl, pl, El, ur, vr, pr, Er;
    prims<M>(lo[0], lo[1], lo[2], lo[3], g1, ul, vl, pl, El);
    prims<M>(hi[0], hi[1], hi[2], hi[3], g1, ur, vr, pr, Er);
    R Fl[4], Fr[4];
    flux<R, X>(lo[0], ul, vl, pl, El, Fl);
    flux<R, X>(hi[0], ur, vr, pr, Er, Fr);
    for (int k = 0; k < 4; k++) {
        const R d = Fl[k] - Fr[k];
        lo[k] += h * d;
        hi[k] += h * d;
    }
}

// calculate_spatial_update (muscl_hancock.cpp:364-371, godunov.cpp:97-111)
// and the forward Euler step (dt*L)+U (forward_euler.cpp:27-28): the
// reference's order in FP64, multiply-adds in FP32.
template <class M>
__device__ __forceinline__ void update_cell(typename M::R rdx, typename M::R rdy, typename M::R dt,
                                            const typename M::R Fw[4], const typename M::R Fe[4],
                                            const typename M::R Gs[4], const typename M::R Gn[4],
                                            const typename M::R U[4], typename M::R L[4], typename M::R Un[4])
{
    for (int k = 0; k < 4; k++) {
        if constexpr (M::fast) {
            L[k] = f32::fma(rdx, Fw[k] - Fe[k], rdy * (Gs[k] - Gn[k]));
            Un[k] = f32::fma(dt, L[k], U[k]);
        } else {
            L[k] = rdx * (Fw[k] - Fe[k]) + rdy * (Gs[k] - Gn[k]);
            Un[k] = dt * L[k] + U[k];
        }
    }
}

// Bits of a non-negative double; their unsigned order is the numeric order,
// which lets wave-speed maxima use 64-bit integer atomics.
__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ __forceinline__ double bitsd(unsigned long long b) { return __longlong_as_double((long long)b); }

// Largest |u|+c and |v|+c of a conserved state (the per-cell terms of
// TimeStep::operator(), timestep.cpp:42-52); NaN terms are skipped by the caller.
template <class M>
__device__ __forceinline__ void wave_speeds(typename M::R U0, typename M::R U1, typename M::R U2,
                                            typename M::R U3, typename M::R gamma, typename M::R g1,
                                            double& Dx, double& Dy)
{
    typedef typename M::R R;
    if constexpr (M::fast) {
        f32::wave_speeds(U0, U1, U2, U3, gamma, g1, Dx, Dy);
        return;
    }
    const R u = M::div(U1, U0);
    const R v = M::div(U2, U0);
    const R p = g1 * (U3 - R(0.5) * U0 * (u * u + v * v));
    const R c = sound<M>(gamma, p, U0);
    Dx = (double)(M::fabs(u) + c);
    Dy = (double)(M::fabs(v) + c);
}

__device__ __forceinline__ double dmax_nan(double acc, double x) { return (x > acc) ? x : acc; }

// Reduce a per-lane maximum over the wave (NaN-free inputs).
__device__ __forceinline__ double wave_max(double x)
{
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(x, off, 64);
        x = (o > x) ? o : x;
    }
    return x;
}

// Reduce a per-lane minimum over the wave (NaN-free inputs).
__device__ __forceinline__ double wave_min(double x)
{
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(x, off, 64);
        x = (o < x) ? o : x;
    }
    return x;
}

// The step and predictor dt of a block from the device partial maxima
// (TimeStep::operator(), timestep.cpp:29-68).  For a block the minimum of
// dx/(|u|+c) over its cells equals dx / max(|u|+c): correctly rounded
// division is monotone in the divisor.  A block whose terms were all NaN has
// max 0 and contributes dx/0 = inf, which the 1e16 start value absorbs.
__device__ __forceinline__ double cfl_eff(double cfl, long long call)
{
    return (call <= 10) ? 0.2 * cfl : cfl;   // timestep.cpp:60-65,72,75
}

// Step dt and speculative predictor dt of a block: every block of the
// problem at U^n.  The reference's compat predictor dt (simulation.cpp:106-110
// with muscl_hancock.cpp:246) sees the blocks before this one at U^{n+1}; the
// fix-up launches below compute it and recompute a block whenever the two
// differ.  In uniform mode (or for Godunov, which has no predictor) the
// predictor dt is the step dt.
template <class R>
__device__ __forceinline__ void step_dts(const HfStepArgs<R>& a, const HfBlockDev& bd, bool muscl, double& dt_step,
                                         double& dt_pred, long long& kstep, long long& kpred)
{
    // Lane l takes blocks first+l, first+l+64, ...: the per-block loads and
    // divisions run side by side instead of as one dependent chain, then the
    // minima are reduced over the wave.  A minimum of non-NaN values does not
    // depend on the order, so the results equal the sequential loop's.  Every
    // lane of the wave calls this (wave-uniform arguments).
    const HfProbDev pr = a.probs[bd.problem];
    kstep = (long long)a.calls[bd.problem] - pr.cps + 1;
    kpred = kstep + 1 + bd.lb;
    double dtx = 1e16, dty = 1e16;
    const int end = pr.first + pr.nb;
    for (int c = pr.first + (int)(threadIdx.x & 63); c < end; c += 64) {
        const double cdx = a.blocks[c].dx, cdy = a.blocks[c].dy;
        const double gx = bitsd(a.gmax[2 * c]), gy = bitsd(a.gmax[2 * c + 1]);
        const double ox = smax(bitsd(a.imax_old[2 * c]), gx), oy = smax(bitsd(a.imax_old[2 * c + 1]), gy);
        dtx = smin(dtx, cdx / ox);
        dty = smin(dty, cdy / oy);
    }
    dtx = wave_min(dtx);
    dty = wave_min(dty);
    dt_step = cfl_eff(a.cfl, kstep) * smin(dtx, dty);
    dt_pred = (muscl && a.dt_mode == 0) ? cfl_eff(a.cfl, kpred) * smin(dtx, dty) : dt_step;
}

// ------------------------------------------------------ fix-up planning --
// The compat predictor dt of block l of a problem is cfl_eff(k_l) times
//     min(1e16, min over c < l of pre_c, min over c >= l of old_c)
// with old_c = min(dx_c / max(old x maxima, x ghosts), dy_c / max(old y
// maxima, y ghosts)) (block c at U^n) and pre_c the same with block c's final
// U^{n+1} maxima: a minimum over one set of non-NaN terms, so any grouping
// gives the reference's bits (step_dts takes the x and y minima apart and
// then their minimum, the same set).  The speculative dt of the main list is
// cfl_eff(k_l) * min(1e16, min over all c of old_c); the step dt is
// cfl_eff(k_step) times the same minimum.

template <class R>
__device__ __forceinline__ double term_old(const HfStepArgs<R>& a, int c)
{
    const double cdx = a.blocks[c].dx, cdy = a.blocks[c].dy;
    const double gx = bitsd(a.gmax[2 * c]), gy = bitsd(a.gmax[2 * c + 1]);
    return smin(cdx / smax(bitsd(a.imax_old[2 * c]), gx), cdy / smax(bitsd(a.imax_old[2 * c + 1]), gy));
}

// pre_c from the given new maxima bit patterns
template <class R>
__device__ __forceinline__ double term_new(const HfStepArgs<R>& a, int c, unsigned long long mx, unsigned long long my)
{
    const double cdx = a.blocks[c].dx, cdy = a.blocks[c].dy;
    const double gx = bitsd(a.gmax[2 * c]), gy = bitsd(a.gmax[2 * c + 1]);
    return smin(cdx / smax(bitsd(mx), gx), cdy / smax(bitsd(my), gy));
}

// Which maxima are final for pre_c:
//  PRE_MAIN  - the main launch's (at the start of the guess launch nothing has
//              been recomputed);
//  PRE_GUESS - the guess launch's where it recomputed the block, else the
//              main launch's (at the start of the check launch);
//  PRE_CHECK - in the check launch, also the blocks it has recomputed so far:
//              written by other waves of the same launch, so the check
//              launch's slot is read with coherent_read.
enum { PRE_MAIN = 0, PRE_GUESS = 1, PRE_CHECK = 2 };

// A word that other waves of the same launch update with device-scope
// atomics, read with a device-scope atomic read-modify-write that changes
// nothing: it is performed where those atomics are, so it sees every one of
// them that has completed.  A plain atomic load may be served by the reading
// XCD's own L2 from a line cached before those updates.
__device__ __forceinline__ unsigned coherent_read(uint32_t* p)
{
    return __hip_atomic_fetch_or(p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long coherent_read(unsigned long long* p)
{
    return __hip_atomic_fetch_or(p, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class R, int SRC>
__device__ __forceinline__ double term_pre(const HfStepArgs<R>& a, int nb_all, int c)
{
    unsigned long long* fix1 = a.imax_fix + 2 * nb_all;
    if (SRC == PRE_CHECK) {
        const bool f2 = coherent_read(&a.fixed[nb_all + c]) != 0u;
        const unsigned long long x2 = coherent_read(&fix1[2 * c]);
        const unsigned long long y2 = coherent_read(&fix1[2 * c + 1]);
        if (f2) return term_new(a, c, x2, y2);
    }
    if (SRC != PRE_MAIN && a.fixed[c]) return term_new(a, c, a.imax_fix[2 * c], a.imax_fix[2 * c + 1]);
    return term_new(a, c, a.imax_new[2 * c], a.imax_new[2 * c + 1]);
}

// Inclusive minimum over lanes 0..lane (scan_lo) or lane..63 (scan_hi).
__device__ __forceinline__ double scan_min_lo(double x)
{
    const int lane = threadIdx.x & 63;
    for (int d = 1; d < 64; d <<= 1) {
        const double o = __shfl_up(x, d, 64);
        if (lane >= d) x = (o < x) ? o : x;
    }
    return x;
}
__device__ __forceinline__ double scan_min_hi(double x)
{
    const int lane = threadIdx.x & 63;
    for (int d = 1; d < 64; d <<= 1) {
        const double o = __shfl_down(x, d, 64);
        if (lane + d < 64) x = (o < x) ? o : x;
    }
    return x;
}

// The fix-up plan of one problem from maxima written by earlier launches
// (plain loads), in one pass over its blocks, 64 at a time (lane <-> block):
// suffix minima of old_c and prefix minima of pre_c give every block's
// predictor dt at once.
//  GUESS (SRC = PRE_MAIN): returns the first block l >= 1 the main list got
//    wrong (skipped, or its speculative dt differs from the exact one) and in
//    m that block's minimum; nb if none.  The guess launch computes every
//    block from there on with cfl_eff(k_l) * min(1e16, m), the exact dt in
//    case no block it recomputes sets a new minimum, and needs to touch
//    block l >= first only if that differs from its speculative dt or the
//    main list skipped it.
//  CHECK (SRC = PRE_GUESS, m = the guess minimum): returns the first block
//    l >= 1 whose computed dt is not the exact one, nb if none, and records
//    the next step's skip verdicts (exact != speculative) of the blocks up to
//    it when `record`.
// O: the problem's speculative minimum min(1e16, min over c of old_c).
template <class R, int SRC>
__device__ int plan_scan(const HfStepArgs<R>& a, const HfProbDev& pr, int nb_all, long long kstep, double O,
                         double& m, bool record)
{
    const int lane = threadIdx.x & 63;
    const int nb = pr.nb;
    const double inf = __builtin_huge_val();
    double carry = inf;   // minimum of pre_c over the chunks before this one
    for (int c0 = 0; c0 < nb; c0 += 64) {
        const int c = c0 + lane;
        const bool in = c < nb;
        const int b = pr.first + (in ? c : 0);
        double o = inf, n = inf;
        if (in) {
            o = term_old(a, b);
            n = term_pre<R, SRC>(a, nb_all, b);
        }
        double tail = inf;   // minimum of old_c after this chunk
        for (int t = c0 + 64 + lane; t < nb; t += 64) tail = smin(tail, term_old(a, pr.first + t));
        tail = wave_min(tail);
        double pre = __shfl_up(scan_min_lo(n), 1, 64);
        pre = (lane == 0 || pre > carry) ? carry : pre;
        double suf = scan_min_hi(o);
        suf = (tail < suf) ? tail : suf;
        const double e = (pre < suf) ? pre : suf;
        const double cf = cfl_eff(a.cfl, kstep + 1 + c);
        const double exact = cf * smin(1e16, e);
        const double spec = cf * smin(1e16, O);
        bool busy;
        if (SRC == PRE_MAIN) {
            busy = in && c >= 1 && (a.skip[b] != 0u || spec != exact);
        } else {
            const bool g = a.fixed[b] != 0u;
            const double used = g ? cf * smin(1e16, m) : spec;
            busy = in && c >= 1 && ((!g && a.skip[b] != 0u) || used != exact);
        }
        const unsigned long long bal = __ballot(busy);
        const int first = bal ? __builtin_ctzll(bal) : 64;
        if (SRC == PRE_GUESS && record && in && c >= 1 && lane <= first) a.skip_next[b] = spec != exact ? 1u : 0u;
        if (bal) {
            if (SRC != PRE_GUESS) m = __shfl(e, first, 64);
            return c0 + first;
        }
        carry = smin(carry, wave_min(n));
    }
    return nb;
}

// The check launch's verdict on block lb of problem p once every block
// before it is final (the launch has completed every earlier round): the
// step dt, the exact predictor dt and whether the block must be (re)computed
// -- its dt so far (the guess launch's, with guess minimum mg, or the main
// list's speculative one) is not the exact one.  Records the next step's skip
// verdict.
template <class R>
__device__ __forceinline__ bool check_verdict(const HfStepArgs<R>& a, int p, const HfProbDev& pr, int nb_all, int lb,
                                              double mg, double& dt_step, double& dt_pred, long long& kstep,
                                              long long& kpred)
{
    const int block = pr.first + lb;
    kstep = (long long)a.calls[p] - pr.cps + 1;
    kpred = kstep + 1 + lb;
    double O = 1e16, P = 1e16;
    const int end = pr.first + pr.nb;
    for (int c = pr.first + (int)(threadIdx.x & 63); c < end; c += 64) {
        const double o = term_old(a, c);
        O = smin(O, o);
        P = smin(P, c < block ? term_pre<R, PRE_CHECK>(a, nb_all, c) : o);
    }
    O = wave_min(O);
    P = wave_min(P);
    const double cf = cfl_eff(a.cfl, kpred);
    dt_step = cfl_eff(a.cfl, kstep) * O;
    dt_pred = cf * P;
    const double spec = cf * O;
    const bool g = a.fixed[block] != 0u;
    const double used = g ? cf * smin(1e16, mg) : spec;
    a.skip_next[block] = spec != dt_pred ? 1u : 0u;   // the same value from every wave
    return (!g && a.skip[block] != 0u) || used != dt_pred;
}

}  // namespace hf
