// hf_uns.hip -- gfx950 kernels of the unstructured quadrilateral-cell
// Godunov solver (src/fvm_2d_uns), FP64, bit-exact with the reference:
// compiled with -ffp-contract=off, every expression in the reference's order.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "hf_common.h"
#include "hf_physics.h"
#include "hf_uns.h"

namespace hfu {

using hf::MathF64;
typedef double R;

__device__ __forceinline__ unsigned long long bits_of(double d) { return (unsigned long long)__double_as_longlong(d); }
__device__ __forceinline__ double double_of(unsigned long long b) { return __longlong_as_double((long long)b); }

// EulerIdealGasModel::rotate_flow_values (model/euler_ideal_gas/euler_ideal_gas.cpp:152-164)
__device__ __forceinline__ void rotate(double c, double s, const double W[4], double Rv[4])
{
    Rv[0] = W[0];
    Rv[1] = c * W[1] + s * W[2];
    Rv[2] = -1.0 * s * W[1] + c * W[2];
    Rv[3] = W[3];
}

// TimeStep::operator()'s per-cell quotient (scheme/timestep/timestep.cpp:41-52)
__device__ __forceinline__ double cell_dt(const double U[4], double dr, double gamma, double g1)
{
    double u, v, p, E;
    hf::prims<MathF64>(U[0], U[1], U[2], U[3], g1, u, v, p, E);
    const double a = hf::sound<MathF64>(gamma, p, U[0]);
    return dr / (__builtin_fabs(__builtin_sqrt(u * u + v * v)) + a);
}

// minimum over the wave of bit patterns (positive doubles order like their bits;
// NaN patterns are larger than every number, as std::min ignores NaN quotients)
__device__ __forceinline__ unsigned long long wave_min_bits(unsigned long long b)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(b, off);
        b = o < b ? o : b;
    }
    return b;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  The grid is padded to a multiple of 8 and workgroup b works on cell
// chunk (b % 8) * (grid / 8) + b / 8, so each XCD sweeps one contiguous
// eighth of the cells and finds most neighbour states in its own L2.
constexpr int XCDS = 8;
__device__ __forceinline__ int cell_chunk()
{
    return (int)(blockIdx.x % XCDS) * (int)(gridDim.x / XCDS) + (int)(blockIdx.x / XCDS);
}
inline dim3 chunk_grid(int n) { return dim3((unsigned)(((n + 255) / 256 + XCDS - 1) / XCDS * XCDS)); }

// The workgroup's minimum into one way of a slot: wave minima through LDS,
// then one atomic.  Called by every thread of the workgroup.
__device__ __forceinline__ void block_min_to_slot(unsigned long long cand, unsigned long long* slot)
{
    __shared__ unsigned long long wmin[4];
    cand = wave_min_bits(cand);
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = cand;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long b = wmin[0];
        for (int w = 1; w < 4; w++) b = wmin[w] < b ? wmin[w] : b;
        atomicMin(slot + (blockIdx.x % SLOT_WAYS) * SLOT_STRIDE, b);
    }
}

// GodunovScheme::operator() (scheme/godunov/godunov.cpp:31-77), boundary
// states of BoundaryConditions (scheme/boundary/boundary.cpp:29-69), and
// ForwardEulerODESolver (scheme/ode/forward_euler/forward_euler.cpp:16-30).
// Four waves per SIMD: the compiler's free choice is 132 VGPRs (three waves);
// at 128 it needs no spills and the latency of the neighbour gathers hides
// better (0.0986 -> 0.0931 ms per step on 1.47 M cells).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_uns_step(const UnsMesh m,
                                                                                          const UnsStep s)
{
    const int c = cell_chunk() * 256 + threadIdx.x;
    const int n = m.n;
    const double gamma = m.gamma, g1 = m.gamma - 1.0;
    const int lane = threadIdx.x & 63;
    double dt;
    const long long kcall = s.kdev ? (long long)s.kdev[s.kpar] + 1 : s.kcall;
    if (s.use_fixed) {
        dt = s.fixed_dt;
    } else {
        const double cf = (kcall <= 10) ? 0.2 * m.cfl : m.cfl;   // timestep.cpp:57-61
        dt = cf * double_of(wave_min_bits(s.slots[s.rd * SLOT_WORDS + lane * SLOT_STRIDE]));
    }
    if (blockIdx.x == 0 && threadIdx.x < 64) {
        if (threadIdx.x == 0 && s.kdev) s.kdev[s.kpar ^ 1] = (unsigned long long)kcall;
        if (threadIdx.x == 0 && !s.use_fixed && kcall >= 1 && kcall <= s.hist_cap) s.hist[kcall - 1] = dt;
        if (s.rs >= 0) s.slots[s.rs * SLOT_WORDS + lane * SLOT_STRIDE] = SLOT_EMPTY;
    }
    unsigned long long cand = ~0ull;
    if (c < n) {
        double U[4];
#pragma unroll
        for (int k = 0; k < 4; k++) U[k] = s.A[(size_t)k * n + c];
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        double shortest = 1e16;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int nb = m.nbr[(size_t)e * n + c];
            const int ct = m.cs_terms;
            const double cr = m.cs[(size_t)(ct * e + 0) * n + c], sr = m.cs[(size_t)(ct * e + 1) * n + c];
            double cb = cr, sb = -sr;   // the back rotation mirrors the forward one (checked on the host)
            if (ct == 4) {
                cb = m.cs[(size_t)(4 * e + 2) * n + c];
                sb = m.cs[(size_t)(4 * e + 3) * n + c];
            }
            double V[4];
            if (nb >= 0) {
#pragma unroll
                for (int k = 0; k < 4; k++) V[k] = s.A[(size_t)k * n + nb];
            } else if (nb == -1) {   // reflective (boundary.cpp:48-56)
                double t[4];
                rotate(cr, sr, U, t);
                t[1] *= -1.0;
                rotate(cb, sb, t, V);
            } else if (nb == -2) {   // transmissive
#pragma unroll
                for (int k = 0; k < 4; k++) V[k] = U[k];
            } else {                 // inlet
#pragma unroll
                for (int k = 0; k < 4; k++) V[k] = m.inlet[k];
            }
            double rU[4], rV[4], F[4], f[4];
            rotate(cr, sr, U, rU);
            rotate(cr, sr, V, rV);
            const hf::State<R> L = hf::make_state<MathF64>(rU, gamma, g1);
            const hf::State<R> Rs = hf::make_state<MathF64>(rV, gamma, g1);
            hf::hllc<MathF64, true>(L, Rs, F);
            rotate(cb, sb, F, f);
            const double len = m.len[(size_t)e * n + c];
            shortest = (shortest < len) ? shortest : len;   // Polygon's incircle radius (polygon.cpp:79-91)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                f[k] *= len;
                acc[k] += f[k];
            }
        }
        const double iv = m.inv_vol[c];
        double N[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            double x = acc[k];
            x *= iv;
            x *= -1.0;
            x *= dt;
            x += U[k];
            N[k] = x;
            s.B[(size_t)k * n + c] = x;
        }
        if (s.wr >= 0) cand = bits_of(cell_dt(N, shortest, gamma, g1));
    }
    if (s.wr >= 0) block_min_to_slot(cand, s.slots + s.wr * SLOT_WORDS);
}

__global__ __launch_bounds__(256) void k_uns_scan(const UnsMesh m, const double* __restrict__ A,
                                                  unsigned long long* slot)
{
    const int c = cell_chunk() * 256 + threadIdx.x;
    const int n = m.n;
    unsigned long long cand = ~0ull;
    if (c < n) {
        double U[4];
#pragma unroll
        for (int k = 0; k < 4; k++) U[k] = A[(size_t)k * n + c];
        double shortest = 1e16;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const double len = m.len[(size_t)e * n + c];
            shortest = (shortest < len) ? shortest : len;
        }
        cand = bits_of(cell_dt(U, shortest, m.gamma, m.gamma - 1.0));
    }
    block_min_to_slot(cand, slot);
}

__global__ void k_uns_reset(unsigned long long* slot) { slot[threadIdx.x * SLOT_STRIDE] = SLOT_EMPTY; }

hipError_t launch_uns_step(const UnsMesh& m, const UnsStep& s, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    if (e0 || e1) hipExtLaunchKernelGGL(k_uns_step, chunk_grid(m.n), dim3(256), 0, st, e0, e1, 0, m, s);
    else hipLaunchKernelGGL(k_uns_step, chunk_grid(m.n), dim3(256), 0, st, m, s);
    return hipGetLastError();
}

hipError_t launch_uns_scan(const UnsMesh& m, const double* A, unsigned long long* slots, int slot, hipStream_t st)
{
    hipLaunchKernelGGL(k_uns_reset, dim3(1), dim3(SLOT_WAYS), 0, st, slots + slot * SLOT_WORDS);
    hipLaunchKernelGGL(k_uns_scan, chunk_grid(m.n), dim3(256), 0, st, m, A, slots + slot * SLOT_WORDS);
    return hipGetLastError();
}

}  // namespace hfu
