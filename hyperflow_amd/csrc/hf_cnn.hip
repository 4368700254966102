// hf_cnn.hip -- gfx950 kernels of the learned coarse-grid correction:
// BF16 MFMA 3x3 convolutions (forward, data gradient, weight gradient) with
// FP32 accumulation, input packing with a shock sensor, the loss, gradient
// reduction, Adam, weight packing, and the state transfers between the
// solver and the network.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hf_cnn.h"

namespace hfn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN -> quiet NaN).
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float(h << 16); }

__device__ __forceinline__ bf16x8 load8(const uint16_t* p)
{
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ bf16x8 zero8()
{
    const uint4 v = make_uint4(0u, 0u, 0u, 0u);
    return __builtin_bit_cast(bf16x8, v);
}

// ------------------------------------------------------------- convolution --
// One wave = one 16-pixel segment of a row, walking `rows` rows.  Lane
// (g = lane/16, j = lane%16) feeds pixel x0+j and K-chunk g of each MFMA and
// receives output channels 16*mt + 4g + r (r = 0..3) of pixel x0+j.
// K-chunk c = 4s+g of MFMA step s is 8 input channels of one tap:
// CIN = 32: tap s, channels 8g..8g+7;  CIN = 8: tap c (c >= 9 is zero).
template <int CIN, int COUT, int MODE>
__global__ __launch_bounds__(256, 3) void k_conv(const ConvArgs a)
{
    constexpr int MT = (COUT + 15) / 16;
    constexpr int S = (9 * CIN / 8 + 3) / 4;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    const int nseg = a.nx >> 4;
    const int seg = wave % nseg;
    const int y0 = (wave / nseg) * a.rows;
    if (y0 >= a.ny) return;
    const int y1 = min(a.ny, y0 + a.rows);
    const int x = seg * 16 + j;
    const int W2 = a.nx + 2;

    bf16x8 wA[MT][S];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int s = 0; s < S; s++) wA[mt][s] = load8(a.wa + ((size_t)(mt * S + s) * 64 + lane) * 8);
    float bias[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int o = 16 * mt + 4 * g + r;
            bias[mt][r] = (MODE != CONV_BWD && o < a.co_real) ? a.bias[o] : 0.0f;
        }

    uint32_t nfb = 0;
    // A row's epilogue: bias, then ReLU (hidden layer), the stored
    // activation's ReLU mask (data gradient), the FP32 output, or the
    // correction applied to the solver state with its fallbacks.
    auto finish = [&](const int y, const f32x4 (&acc)[MT]) {
        const size_t pix = (size_t)(y + 1) * W2 + (x + 1);
        if (MODE == CONV_HIDDEN || MODE == CONV_BWD) {
#pragma unroll
            for (int mt = 0; mt < MT; mt++) {
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] = acc[mt][r] + bias[mt][r];
                uint16_t* dst = a.out + pix * COUT + 16 * mt + 4 * g;
                if (MODE == CONV_HIDDEN) {
#pragma unroll
                    for (int r = 0; r < 4; r++) v[r] = v[r] > 0.0f ? v[r] : 0.0f;
                } else {
                    const uint2 m = *reinterpret_cast<const uint2*>(a.mask + pix * COUT + 16 * mt + 4 * g);
                    const uint32_t mv[4] = {m.x & 0xffffu, m.x >> 16, m.y & 0xffffu, m.y >> 16};
#pragma unroll
                    for (int r = 0; r < 4; r++) v[r] = (mv[r] != 0u && mv[r] < 0x8000u) ? v[r] : 0.0f;
                }
                uint2 pk;
                pk.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
                pk.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
                *reinterpret_cast<uint2*>(dst) = pk;
            }
        } else if (g == 0) {
            // output channels 0..3 of pixel x (lanes 0..15)
            float v[4];
#pragma unroll
            for (int r = 0; r < 4; r++) v[r] = acc[0][r] + bias[0][r];
            if (MODE == CONV_OUT) {
                const size_t np = (size_t)a.nx * a.ny;
#pragma unroll
                for (int r = 0; r < 4; r++) a.outf[r * np + (size_t)y * a.nx + x] = v[r];
            } else {
                float* U = a.U + (size_t)y * a.pitch + x;
                float u[4], n[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    u[r] = U[r * a.plane];
                    n[r] = u[r] + a.scale[r] * v[r];
                }
                const float p = (a.gamma - 1.0f) * (n[3] - 0.5f * (n[1] * n[1] + n[2] * n[2]) / n[0]);
                const bool shock = a.shock && a.shock[(size_t)y * a.nx + x];
                bool inrange = true;
                if (a.max_output > 0.0f)
#pragma unroll
                    for (int r = 0; r < 4; r++) inrange = inrange && fabsf(v[r]) <= a.max_output;
                const bool ok = n[0] > 0.0f && p > 0.0f && !shock && inrange && isfinite(p) && isfinite(n[1]) &&
                                isfinite(n[2]) && isfinite(n[3]);
                if (ok) {
#pragma unroll
                    for (int r = 0; r < 4; r++) U[r * a.plane] = n[r];
                } else {
                    nfb++;
                }
            }
        }
    };

    if (CIN == 32) {
        // The B operand of tap (dy, dx) at output row y is padded input row
        // y+dy+1 at column x+dx: each output row needs three input rows of
        // three column shifts, and each new row costs three loads.  Input
        // rows live in a ring of RING buffers named at compile time (the row
        // loop is unrolled by RING, so the window slides by renaming, not by
        // copying registers, which would wait for the loads): the loads of
        // row y+2+PF are issued while row y is computed, PF rows ahead of
        // their use, so the memory latency overlaps PF rows of MFMAs.
        constexpr int PF = 2, RING = 3 + PF;
        bf16x8 Rb[RING][3];
        auto load_row = [&](const int k, bf16x8 (&dst)[3]) {   // padded row y0 + k
            const int yr = min(y0 + k, a.ny + 1);
#pragma unroll
            for (int d = 0; d < 3; d++) dst[d] = load8(a.in + ((size_t)yr * W2 + (x + d)) * CIN + g * 8);
        };
#pragma unroll
        for (int k = 0; k < RING - 1; k++) load_row(k, Rb[k]);
        for (int base = y0; base < y1; base += RING) {
#pragma unroll
            for (int u = 0; u < RING; u++) {
                const int y = base + u;
                if (y < y1) {
                    load_row(y - y0 + RING - 1, Rb[(u + RING - 1) % RING]);
                    f32x4 acc[MT];
#pragma unroll
                    for (int mt = 0; mt < MT; mt++) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int t = 0; t < 9; t++)
#pragma unroll
                        for (int mt = 0; mt < MT; mt++)
                            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wA[mt][t], Rb[(u + t / 3) % RING][t % 3],
                                                                             acc[mt], 0, 0, 0);
                    finish(y, acc);
                }
            }
        }
    } else {
        for (int y = y0; y < y1; y++) {
            f32x4 acc[MT];
#pragma unroll
            for (int mt = 0; mt < MT; mt++) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < S; s++) {
                int tap = 4 * s + g;
                const bool valid = tap < 9;
                tap = valid ? tap : 8;
                const int dy = tap / 3 - 1, dx = tap % 3 - 1;
                bf16x8 B = load8(a.in + ((size_t)(y + 1 + dy) * W2 + (x + 1 + dx)) * CIN);
                B = valid ? B : zero8();
#pragma unroll
                for (int mt = 0; mt < MT; mt++)
                    acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wA[mt][s], B, acc[mt], 0, 0, 0);
            }
            finish(y, acc);
        }
    }
    if (MODE == CONV_APPLY) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) nfb += __shfl_xor(nfb, off);
        if (lane == 0 && nfb) atomicAdd(a.fallback, nfb);
    }
}

// ---------------------------------------------------------- weight gradient --
// dW[tap][ci][co] = sum_p a[p + tap][ci] * gz[p][co],  db[co] = sum_p gz[p][co].
// A workgroup owns a 64-pixel column strip over `rows` rows.  Per output row
// it stages gz (channel-major, gT[co][px]) and one new input row in LDS,
// channel-major and shifted by each dx (aT[slot][dx][ci][px] =
// a[row][x0+px+dx][ci]); the three input rows a tap can reach live in a ring
// of three slots, so each input row is staged once.  Each lane's 8
// consecutive pixels of one channel are then one aligned 16-byte LDS read.
// The pixels are the MFMA reduction dimension: A = gz^T (rows = co), B =
// shifted input (columns = ci).  Each wave owns a fixed set of (tap, co
// tile, ci tile) products; partial sums go to part[workgroup].
constexpr int LPX = 72;   // LDS row pitch in elements (64 pixels + 8 padding)

template <int CI, int CO>
__global__ __launch_bounds__(256) void k_wgrad(const WgradArgs a)
{
    constexpr int CIL = CI < 16 ? 16 : CI;
    constexpr int COL = CO < 16 ? 16 : CO;
    constexpr int MT = COL / 16, NT = CIL / 16;
    constexpr int ITEMS = 9 * MT * NT;
    constexpr int NPW = (ITEMS + 3) / 4;
    __shared__ __attribute__((aligned(16))) uint16_t aT[3][3][CIL][LPX];
    __shared__ __attribute__((aligned(16))) uint16_t gT[COL][LPX];

    const int tid = threadIdx.x;
    const int lane = tid & 63, g = lane >> 4, j = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nstrip = a.nx >> 6;
    const int x0 = (blockIdx.x % nstrip) * 64;
    const int y0 = (blockIdx.x / nstrip) * a.rows;
    const int y1 = min(a.ny, y0 + a.rows);
    const int W2 = a.nx + 2;

    // zero the padding channels once (never written below)
    if (CIL > CI)
        for (int i = tid; i < 9 * (CIL - CI) * LPX; i += 256) {
            const int t = i / ((CIL - CI) * LPX), rem = i % ((CIL - CI) * LPX);
            aT[t / 3][t % 3][CI + rem / LPX][rem % LPX] = 0;
        }
    if (COL > CO)
        for (int i = tid; i < (COL - CO) * LPX; i += 256) gT[CO + i / LPX][i % LPX] = 0;

    // input row r (interior index, -1 <= r <= ny) into ring slot (r + 3) % 3
    auto stage_row = [&](int r) {
        const int slot = (r + 3) % 3;
        for (int i = tid; i < 66 * (CI / 8); i += 256) {
            const int ps = i / (CI / 8) - 1, cb = i % (CI / 8);
            const uint4 v = *reinterpret_cast<const uint4*>(a.a + ((size_t)(r + 1) * W2 + x0 + ps + 1) * CI + cb * 8);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int dxi = 0; dxi < 3; dxi++) {
                const int px = ps - (dxi - 1);
                if (px >= 0 && px < 64) {
#pragma unroll
                    for (int e = 0; e < 8; e++)
                        aT[slot][dxi][cb * 8 + e][px] = (uint16_t)(w[e >> 1] >> (16 * (e & 1)));
                }
            }
        }
    };

    f32x4 acc[NPW];
#pragma unroll
    for (int u = 0; u < NPW; u++) acc[u] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float db = 0.0f;

    stage_row(y0 - 1);
    stage_row(y0);
    for (int y = y0; y < y1; y++) {
        __syncthreads();
        for (int i = tid; i < 64 * (CO / 8); i += 256) {
            const int px = i / (CO / 8), cb = i % (CO / 8);
            const uint4 v = *reinterpret_cast<const uint4*>(a.gz + ((size_t)(y + 1) * W2 + x0 + px + 1) * CO + cb * 8);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int e = 0; e < 8; e++) gT[cb * 8 + e][px] = (uint16_t)(w[e >> 1] >> (16 * (e & 1)));
        }
        stage_row(y + 1);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NPW; u++) {
            const int q = wave + 4 * u;
            if (q < ITEMS) {
                const int tap = q % 9, mt = (q / 9) % MT, nt = (q / 9) / MT;
                const int slot = (y + tap / 3 + 2) % 3;   // input row y + dy, dy = tap/3 - 1
#pragma unroll
                for (int ks = 0; ks < 2; ks++) {
                    const bf16x8 A = load8(&gT[16 * mt + j][32 * ks + 8 * g]);
                    const bf16x8 B = load8(&aT[slot][tap % 3][16 * nt + j][32 * ks + 8 * g]);
                    acc[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[u], 0, 0, 0);
                }
            }
        }
        if (tid < CO)
            for (int px = 0; px < 64; px++) db += bf2f(gT[tid][px]);
    }

    float* P = a.part + (size_t)blockIdx.x * (9 * CI * CO + CO);
#pragma unroll
    for (int u = 0; u < NPW; u++) {
        const int q = wave + 4 * u;
        if (q < ITEMS) {
            const int tap = q % 9, mt = (q / 9) % MT, nt = (q / 9) / MT;
            const int ci = 16 * nt + j;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int co = 16 * mt + 4 * g + r;
                if (co < CO && ci < CI) P[((size_t)tap * CI + ci) * CO + co] = acc[u][r];
            }
        }
    }
    if (tid < CO) P[9 * CI * CO + tid] = db;
}

// ------------------------------------------------------------------ helpers --
__device__ __forceinline__ float pressure(const float* U, int64_t plane, float gamma)
{
    const float r = U[0], mu = U[plane], mv = U[2 * plane], E = U[3 * plane];
    return (gamma - 1.0f) * (E - 0.5f * (mu * mu + mv * mv) / r);
}

__global__ __launch_bounds__(256) void k_pack(const PackArgs a)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= a.nx) return;
    const float* U = a.U + (size_t)y * a.pitch + x;
    float v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) v[r] = (U[r * a.plane] - a.mean[r]) * a.iscale[r];
    uint4 pk;
    pk.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
    pk.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
    pk.z = 0u;
    pk.w = 0u;
    *reinterpret_cast<uint4*>(a.X + ((size_t)(y + 1) * (a.nx + 2) + x + 1) * 8) = pk;
    if (a.shock) {
        // normalised second difference of the pressure, the larger of x and y
        const int xl = x > 0 ? -1 : 0, xr = x < a.nx - 1 ? 1 : 0;
        const int yd = y > 0 ? -a.pitch : 0, yu = y < a.ny - 1 ? a.pitch : 0;
        const float p = pressure(U, a.plane, a.gamma);
        const float pl = pressure(U + xl, a.plane, a.gamma), pr = pressure(U + xr, a.plane, a.gamma);
        const float pd = pressure(U + yd, a.plane, a.gamma), pu = pressure(U + yu, a.plane, a.gamma);
        const float sx = fabsf(pr - 2.0f * p + pl) / (pr + 2.0f * p + pl);
        const float sy = fabsf(pu - 2.0f * p + pd) / (pu + 2.0f * p + pd);
        a.shock[(size_t)y * a.nx + x] = (sx > a.theta || sy > a.theta || !(sx == sx) || !(sy == sy)) ? 1 : 0;
    }
}

__global__ __launch_bounds__(256) void k_loss(const LossArgs a)
{
    __shared__ float red[256];
    const int64_t np = (int64_t)a.nx * a.ny;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    if (i < np) {
        const int y = (int)(i / a.nx), x = (int)(i % a.nx);
        float gv[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float t = (a.T[r * np + i] - a.Uc[r * np + i]) * a.inv_scale[r];
            const float d = a.Y[r * np + i] - t;
            l += d * d;
            gv[r] = a.gscale * 2.0f * d;
        }
        uint4 pk;
        pk.x = (uint32_t)f2bf(gv[0]) | ((uint32_t)f2bf(gv[1]) << 16);
        pk.y = (uint32_t)f2bf(gv[2]) | ((uint32_t)f2bf(gv[3]) << 16);
        pk.z = 0u;
        pk.w = 0u;
        *reinterpret_cast<uint4*>(a.G + ((size_t)(y + 1) * (a.nx + 2) + x + 1) * 8) = pk;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.part[blockIdx.x] = red[0];
}

// Sum of the workgroup partials in two deterministic stages.  Stage 1: block
// (bx, by) adds partials [by*chunk, (by+1)*chunk) of 64 consecutive entries,
// four interleaved subsets per entry combined in a fixed order, into
// tmp[by].  Stage 2 adds the stage-1 rows in order and writes the layer's
// real layout W[9][ci_r][co_r] followed by b[co_r] (padding channels
// dropped).
__global__ __launch_bounds__(256) void k_reduce_stage1(const float* __restrict__ part, int nparts, int stride,
                                                       int chunk, float* __restrict__ tmp)
{
    __shared__ float red[4][64];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int s = blockIdx.x * 64 + tx;
    const int w0 = blockIdx.y * chunk, w1 = min(nparts, w0 + chunk);
    float acc = 0.0f;
    if (s < stride)
        for (int w = w0 + ty; w < w1; w += 4) acc += part[(size_t)w * stride + s];
    red[ty][tx] = acc;
    __syncthreads();
    if (ty == 0 && s < stride) tmp[(size_t)blockIdx.y * stride + s] = (red[0][tx] + red[1][tx]) + (red[2][tx] + red[3][tx]);
}

__global__ __launch_bounds__(256) void k_reduce_grad(const float* __restrict__ part, int nparts, int CI, int CO,
                                                     int ci_r, int co_r, float* __restrict__ gout)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int stride = 9 * CI * CO + CO;
    if (s >= stride) return;
    float v = 0.0f;
    for (int w = 0; w < nparts; w++) v += part[(size_t)w * stride + s];
    int dst = -1;
    if (s < 9 * CI * CO) {
        const int tap = s / (CI * CO), rem = s % (CI * CO);
        const int ci = rem / CO, co = rem % CO;
        if (ci < ci_r && co < co_r) dst = (tap * ci_r + ci) * co_r + co;
    } else {
        const int co = s - 9 * CI * CO;
        if (co < co_r) dst = 9 * ci_r * co_r + co;
    }
    if (dst >= 0) gout[dst] = v;
}

__global__ __launch_bounds__(256) void k_reduce_sum(const float* __restrict__ part, int n, float scale,
                                                    float* __restrict__ out)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * scale;
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                                              float c1, float c2)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_pack_w(const float* __restrict__ W, int ci_r, int co_r, int transposed,
                                                uint16_t* __restrict__ wa)
{
    constexpr int MT = (COUT + 15) / 16;
    constexpr int S = (9 * CIN / 8 + 3) / 4;
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= MT * S * 64 * 8) return;
    const int e = idx & 7, lane = (idx >> 3) & 63;
    const int s = (idx >> 9) % S, mt = (idx >> 9) / S;
    const int o = 16 * mt + (lane & 15), g = lane >> 4;
    int tap, i;
    if (CIN == 32) {
        tap = s;
        i = 8 * g + e;
    } else {
        tap = 4 * s + g;
        i = e;
    }
    float val = 0.0f;
    if (tap < 9) {
        if (!transposed) {
            if (o < co_r && i < ci_r) val = W[((size_t)tap * ci_r + i) * co_r + o];
        } else {
            if (o < ci_r && i < co_r) val = W[((size_t)(8 - tap) * ci_r + o) * co_r + i];
        }
    }
    wa[idx] = f2bf(val);
}

__global__ __launch_bounds__(256) void k_state_copy(const float* __restrict__ src, int64_t splane, int spitch,
                                                    float* __restrict__ dst, int64_t dplane, int dpitch, int nx)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= nx) return;
#pragma unroll
    for (int r = 0; r < 4; r++) dst[r * dplane + (size_t)y * dpitch + x] = src[r * splane + (size_t)y * spitch + x];
}

__global__ __launch_bounds__(256) void k_restrict2(const float* __restrict__ U, int64_t plane, int pitch, int nxc,
                                                   int nyc, float* __restrict__ dst)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= nxc) return;
    const size_t np = (size_t)nxc * nyc;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const float* u = U + r * plane + (size_t)(2 * y) * pitch + 2 * x;
        dst[r * np + (size_t)y * nxc + x] = 0.25f * ((u[0] + u[1]) + (u[pitch] + u[pitch + 1]));
    }
}

// ---------------------------------------------------------------- launchers --
namespace {
template <int CIN, int COUT, int MODE>
hipError_t conv_t(const ConvArgs& a, hipStream_t s)
{
    const int waves = (a.nx / 16) * ((a.ny + a.rows - 1) / a.rows);
    hipLaunchKernelGGL((k_conv<CIN, COUT, MODE>), dim3((waves + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_conv(int cin, int cout, int mode, const ConvArgs& a, hipStream_t s)
{
    if (cin == 8 && cout == 32 && mode == CONV_HIDDEN) return conv_t<8, 32, CONV_HIDDEN>(a, s);
    if (cin == 32 && cout == 32 && mode == CONV_HIDDEN) return conv_t<32, 32, CONV_HIDDEN>(a, s);
    if (cin == 32 && cout == 8 && mode == CONV_OUT) return conv_t<32, 8, CONV_OUT>(a, s);
    if (cin == 32 && cout == 8 && mode == CONV_APPLY) return conv_t<32, 8, CONV_APPLY>(a, s);
    if (cin == 32 && cout == 32 && mode == CONV_BWD) return conv_t<32, 32, CONV_BWD>(a, s);
    if (cin == 8 && cout == 32 && mode == CONV_BWD) return conv_t<8, 32, CONV_BWD>(a, s);
    return hipErrorInvalidValue;
}

int wgrad_workgroups(int nx, int ny, int rows) { return (nx / 64) * ((ny + rows - 1) / rows); }

hipError_t launch_wgrad(int ci, int co, const WgradArgs& a, int* nwg, hipStream_t s)
{
    const int n = wgrad_workgroups(a.nx, a.ny, a.rows);
    *nwg = n;
    if (ci == 8 && co == 32) hipLaunchKernelGGL((k_wgrad<8, 32>), dim3(n), dim3(256), 0, s, a);
    else if (ci == 32 && co == 32) hipLaunchKernelGGL((k_wgrad<32, 32>), dim3(n), dim3(256), 0, s, a);
    else if (ci == 32 && co == 8) hipLaunchKernelGGL((k_wgrad<32, 8>), dim3(n), dim3(256), 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_pack(const PackArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack, dim3((a.nx + 255) / 256, a.ny), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_loss(const LossArgs& a, int* nblocks, hipStream_t s)
{
    const int64_t np = (int64_t)a.nx * a.ny;
    const int n = (int)((np + 255) / 256);
    *nblocks = n;
    hipLaunchKernelGGL(k_loss, dim3(n), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_reduce_grad(const float* part, int nparts, int ci, int co, int ci_r, int co_r, float* g,
                              float* tmp, hipStream_t s)
{
    const int stride = 9 * ci * co + co;
    const int chunk = (nparts + REDUCE_SPLIT - 1) / REDUCE_SPLIT;
    hipLaunchKernelGGL(k_reduce_stage1, dim3((stride + 63) / 64, REDUCE_SPLIT), dim3(256), 0, s, part, nparts, stride,
                       chunk, tmp);
    hipLaunchKernelGGL(k_reduce_grad, dim3((stride + 255) / 256), dim3(256), 0, s, tmp, REDUCE_SPLIT, ci, co, ci_r,
                       co_r, g);
    return hipGetLastError();
}

hipError_t launch_reduce_sum(const float* part, int n, float scale, float* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(256), 0, s, part, n, scale, out);
    return hipGetLastError();
}

hipError_t launch_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

int packed_weight_elems(int cin, int cout) { return ((cout + 15) / 16) * ((9 * cin / 8 + 3) / 4) * 64 * 8; }

hipError_t launch_pack_weights(int cin, int cout, const float* W, int ci_r, int co_r, int transposed, uint16_t* wa,
                               hipStream_t s)
{
    const int n = packed_weight_elems(cin, cout);
    const dim3 grid((n + 255) / 256);
    if (cin == 8 && cout == 32) hipLaunchKernelGGL((k_pack_w<8, 32>), grid, dim3(256), 0, s, W, ci_r, co_r, transposed, wa);
    else if (cin == 32 && cout == 32)
        hipLaunchKernelGGL((k_pack_w<32, 32>), grid, dim3(256), 0, s, W, ci_r, co_r, transposed, wa);
    else if (cin == 32 && cout == 8)
        hipLaunchKernelGGL((k_pack_w<32, 8>), grid, dim3(256), 0, s, W, ci_r, co_r, transposed, wa);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_state_copy(const float* src, int64_t splane, int spitch, float* dst, int64_t dplane, int dpitch,
                             int nx, int ny, hipStream_t s)
{
    hipLaunchKernelGGL(k_state_copy, dim3((nx + 255) / 256, ny), dim3(256), 0, s, src, splane, spitch, dst, dplane,
                       dpitch, nx);
    return hipGetLastError();
}

hipError_t launch_restrict2(const float* U, int64_t plane, int pitch, int nxc, int nyc, float* dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_restrict2, dim3((nxc + 255) / 256, nyc), dim3(256), 0, s, U, plane, pitch, nxc, nyc, dst);
    return hipGetLastError();
}

}  // namespace hfn
