// shwfs.hip -- MI355X (gfx950) Shack-Hartmann wavefront-sensor hot path and its C ABI.
//
// Forward model of one subaperture (SH.trigger_code, processing_objects/sh.py:364-420):
//   E      = A * exp(i*(2 pi/lambda)*phi) * tilt                        (np x np, zero padded to N)
//   I      = |FFT_N(E)|^2 ;  NGS: psf = fftshift(I);  LGS: psf = Re IFFT_N(FFT_N(I) * K)
//   B      = rebin(cut(mask * psf))                                       (npx x npx CCD pixels)
// Because E has only np x np non-zero samples, |FFT_N(E)|^2 is a trigonometric polynomial whose
// autocorrelation support is (2np-1)^2 lags.  Its power spectrum S on any grid M >= 2np-1
// determines it exactly, and every step after |.|^2 (shift, mask, cut, rebin, and for LGS the
// kernel correlation) is linear.  So the kernels compute
//   NGS:  B = P . S . Q^T            (P, Q: npx x M real binning operators built on the host)
//   LGS:  B = Re(P . (N^2 C(-m) K(m)) . Q^T)   (C: autocorrelation on the lag grid)
// where S = |FFT_M(E)|^2 comes from register-resident pruned FFTs of size M = 16 or 32 instead
// of the N = 64/128 transforms of the reference.  Normalisation to S0*masked_area
// (sh.py:422-428) is exact: either analytic by Parseval (no cut, all-ones mask, so the image sum
// is norm_const * sum(A^2)) or by an FP64 reduction of the image followed by a rescale.
//
// Build: hipcc --offload-arch=gfx950 -O3 -shared -fPIC (see specula_amd/build.py).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/shwfs.h"
#include "fft_reg.h"

using namespace shwfs;

// ============================================================================ error plumbing
namespace {
thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
}  // namespace

#define SHWFS_HIP_CHECK(expr)                                                              \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_error(SHWFS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define SHWFS_LAUNCH_CHECK()                                                               \
    do {                                                                                   \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess)                                                              \
            return set_error(SHWFS_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

// ============================================================================ device helpers
namespace {

constexpr int kWave = 64;
constexpr double kTwoPi = 6.283185307179586476925286766559;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Block-wide FP64 sum; result valid in every thread.  `red` needs blockDim/64 doubles.
__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const int nw = (blockDim.x + kWave - 1) / kWave;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[w];   // fixed order: deterministic
    return s;
}

// Phase of one pupil sample: (2 pi / lambda) * phi - (pi / N) * (ax + ay), reduced to
// [-pi, pi] in FP64 before the FP32 sincos (the unreduced phase reaches tens of radians).
__device__ __forceinline__ float2 field_sample(float a, float ph, double k, double tilt, double ax_ay) {
    double psi = (double)ph * k - tilt * ax_ay;
    psi -= kTwoPi * rint(psi * (1.0 / kTwoPi));
    float s, c;
    sincosf((float)psi, &s, &c);
    return make_float2(a * c, a * s);
}

// Same sample with the reduction done in revolutions: r = psi/(2 pi) - rint(psi/(2 pi)) in FP64,
// then the hardware v_sin_f32 / v_cos_f32 (which take revolutions) on |r| <= 1/2.
__device__ __forceinline__ float2 field_sample_rev(float a, float ph, double k_rev, double tilt_rev, double ax_ay) {
    double rv = (double)ph * k_rev - tilt_rev * ax_ay;
    rv -= rint(rv);
    const float rf = (float)rv;
    return make_float2(a * __builtin_amdgcn_cosf(rf), a * __builtin_amdgcn_sinf(rf));
}

// ---------------------------------------------------------------- amplitude statistics
// Per frame: sum(A) (masked area, electric_field.py:87-89) and sum(A^2), FP64, deterministic.
// Last-block-done pattern: every block writes a partial, the last one to arrive reduces the
// partials in fixed order and resets the arrival counter (so graph replays need no memset).
__global__ void amp_stats_kernel(const float* __restrict__ A, int64_t a_stride, int64_t n_px,
                                 double* __restrict__ part, unsigned* __restrict__ counter,
                                 double* __restrict__ out) {
    __shared__ double red[16];
    __shared__ bool last;
    const int f = blockIdx.y;
    const float* a = A + f * a_stride;
    double s1 = 0.0, s2 = 0.0;
    const int64_t n4 = n_px / 4;
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const bool aligned = (reinterpret_cast<uintptr_t>(a) % 16) == 0;
    if (aligned) {
        for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n4;
             q += (int64_t)gridDim.x * blockDim.x) {
            float4 v = a4[q];
            s1 += (double)v.x + (double)v.y + (double)v.z + (double)v.w;
            s2 += (double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z + (double)v.w * v.w;
        }
        for (int64_t q = 4 * n4 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n_px;
             q += (int64_t)gridDim.x * blockDim.x) {
            double v = a[q];
            s1 += v; s2 += v * v;
        }
    } else {
        for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n_px;
             q += (int64_t)gridDim.x * blockDim.x) {
            double v = a[q];
            s1 += v; s2 += v * v;
        }
    }
    s1 = block_sum(s1, red);
    s2 = block_sum(s2, red);
    double* p = part + ((int64_t)f * gridDim.x + blockIdx.x) * 2;
    if (threadIdx.x == 0) {
        p[0] = s1;
        p[1] = s2;
        __threadfence();
        unsigned prev = atomicAdd(&counter[f], 1u);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (last) {
        __threadfence();
        double t1 = 0.0, t2 = 0.0;
        const double* pf = part + (int64_t)f * gridDim.x * 2;
        for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
            t1 += pf[2 * b];
            t2 += pf[2 * b + 1];
        }
        // per-thread strided partials, then fixed-order block reduction
        t1 = block_sum(t1, red);
        t2 = block_sum(t2, red);
        if (threadIdx.x == 0) {
            out[2 * f] = t1;
            out[2 * f + 1] = t2;
            counter[f] = 0u;
        }
    }
}

// ---------------------------------------------------------------- NGS forward
struct NgsArgs {
    const float* A;
    int64_t a_stride;
    const float* phase;
    int64_t ph_stride;
    float* ccd;
    int64_t ccd_stride;
    const float* roww;   // npx_pad x M   (P)
    const float* colw;   // npx_pad x M   (Q)
    const double* amp;   // [frame][2]: sum(A), sum(A^2)     (frame index f * amp_fstep)
    int amp_fstep;
    double* bpart;       // [frame][n_tiles]: sum of unscaled B   (non-Parseval mode)
    const float2* kspec; // LGS: [n_sub][L][L] kernel spectra on the lag grid (L = 2 np - 1)
    int n_lenses, ef_size, npx, npx_pad, n_sub, n_tiles, parseval;
    double k, tilt, a0, S0, pixel_area, norm_const;
};

// Work split of one block = TS subapertures, TS*NP threads:
//   row pass    : thread (s, y)  -> pruned FFT_M of pupil row y of subaperture s
//   column pass : thread (s, kx) -> pruned FFT_M of column kx, |.|^2, U = P.S  (in registers)
//   bin pass    : thread (s, Y)  -> CCD row Y of subaperture s: B[Y,:] = U[Y,:].Q^T
// LDS holds G1 (complex, TS x M x NP) and later U (real, TS x npx_pad x M) in the same buffer.
template <int NP, int M, int TS>
__global__ void __launch_bounds__(TS * NP)
sh_ngs_forward_kernel(NgsArgs g) {
    extern __shared__ float4 lds_raw[];
    float2* g1 = reinterpret_cast<float2*>(lds_raw);
    float* us = reinterpret_cast<float*>(lds_raw);
    __shared__ double red[8];

    constexpr int G1S = M * NP + 1;            // complex stride per subaperture (odd: spreads banks)
    const int US = g.npx_pad * M + 4;          // float stride per subaperture in U
    constexpr int NT = TS * NP;
    constexpr int CPT = (TS * M + NT - 1) / NT;   // columns per thread

    const int t = threadIdx.x;
    const int f = blockIdx.y;
    const int tile = blockIdx.x;
    const int n = g.n_lenses;

    // ---- row pass
    {
        const int s = t % TS, y = t / TS;
        const int sub = tile * TS + s;
        float2 e[NP];
        if (sub < g.n_sub) {
            const int i = sub / n, j = sub - (sub / n) * n;
            const int64_t off = (int64_t)(i * NP + y) * g.ef_size + (int64_t)j * NP;
            const float* ar = g.A + f * g.a_stride + off;
            const float* pr = g.phase + f * g.ph_stride + off;
            const double ay = g.a0 + y;
#pragma unroll
            for (int x = 0; x < NP; ++x)
                e[x] = field_sample(ar[x], pr[x], g.k, g.tilt, g.a0 + x + ay);
        } else {
#pragma unroll
            for (int x = 0; x < NP; ++x) e[x] = make_float2(0.f, 0.f);
        }
        float2 r[M];
        fft_pruned<M, NP, false>(e, r);
#pragma unroll
        for (int kx = 0; kx < M; ++kx) g1[s * G1S + kx * NP + y] = r[kx];
    }
    __syncthreads();

    // ---- column pass: S = |FFT_M over y|^2, kept in registers across the LDS reuse
    float S[CPT][M];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int c = t + q * NT;
        if (c < TS * M) {
            const int s = c % TS, kx = c / TS;
            float2 col[NP];
#pragma unroll
            for (int y = 0; y < NP; ++y) col[y] = g1[s * G1S + kx * NP + y];
            float2 G[M];
            fft_pruned<M, NP, false>(col, G);
#pragma unroll
            for (int ky = 0; ky < M; ++ky) S[q][ky] = G[ky].x * G[ky].x + G[ky].y * G[ky].y;
        }
    }
    __syncthreads();

    // U[s][Y][kx] = sum_ky P[Y][ky] S[ky][kx]   (P row index uniform across the block)
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int c = t + q * NT;
        if (c < TS * M) {
            const int s = c % TS, kx = c / TS;
            for (int Y0 = 0; Y0 < g.npx_pad; Y0 += 4) {
                float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f, acc3 = 0.f;
                const float* p0 = g.roww + Y0 * M;
#pragma unroll
                for (int ky = 0; ky < M; ++ky) {
                    acc0 = fmaf(p0[ky], S[q][ky], acc0);
                    acc1 = fmaf(p0[M + ky], S[q][ky], acc1);
                    acc2 = fmaf(p0[2 * M + ky], S[q][ky], acc2);
                    acc3 = fmaf(p0[3 * M + ky], S[q][ky], acc3);
                }
                float* u = us + s * US + Y0 * M + kx;
                u[0] = acc0;
                u[M] = acc1;
                u[2 * M] = acc2;
                u[3 * M] = acc3;
            }
        }
    }
    __syncthreads();

    // ---- bin pass: B[Y][X] = sum_kx U[Y][kx] Q[X][kx]; store CCD row segments
    double scale = 1.0;
    if (g.parseval) {
        const double* am = g.amp + (int64_t)f * g.amp_fstep;
        scale = g.S0 * g.pixel_area * am[0] / (g.norm_const * am[1]);
    }
    const int W = n * g.npx;
    double bsum = 0.0;
    for (int c = t; c < TS * g.npx; c += NT) {
        const int s = c % TS, Y = c / TS;
        const int sub = tile * TS + s;
        if (sub >= g.n_sub) continue;
        const int i = sub / n, j = sub - (sub / n) * n;
        float u[M];
#pragma unroll
        for (int kx = 0; kx < M; ++kx) u[kx] = us[s * US + Y * M + kx];
        float* dst = g.ccd + f * g.ccd_stride + (int64_t)(i * g.npx + Y) * W + (int64_t)j * g.npx;
        for (int X0 = 0; X0 < g.npx; X0 += 4) {
            float b[4] = {0.f, 0.f, 0.f, 0.f};
            const float* q0 = g.colw + X0 * M;
#pragma unroll
            for (int kx = 0; kx < M; ++kx) {
                b[0] = fmaf(q0[kx], u[kx], b[0]);
                b[1] = fmaf(q0[M + kx], u[kx], b[1]);
                b[2] = fmaf(q0[2 * M + kx], u[kx], b[2]);
                b[3] = fmaf(q0[3 * M + kx], u[kx], b[3]);
            }
            const int nx = min(4, g.npx - X0);
            for (int q = 0; q < nx; ++q) {
                bsum += (double)b[q];
                dst[X0 + q] = (float)((double)b[q] * scale);
            }
        }
    }
    if (!g.parseval) {
        bsum = block_sum(bsum, red);
        if (t == 0) g.bpart[(int64_t)f * g.n_tiles + tile] = bsum;
    }
}

// ---------------------------------------------------------------- NGS forward, MFMA binning (M = 16)
// One wavefront owns SW = 64/NP consecutive subapertures (a block holds 4 independent waves):
//   A: lane (q, x)  -> E along column x of subaperture q, pruned FFT_16 over y    -> G1[q][ky][x]
//   B: lane (q, ky) -> pruned FFT_16 over x of row ky, |.|^2                      -> S[q][ky][kx]
//   C: per subaperture, B = P S Q^T as 2 x 4 v_mfma_f32_16x16x4f32:
//        T = S Q^T  with the K index permuted as k = 4*(lane/16) + step, so that each lane reads
//                   one float4 of S and one float4 of Q;
//        B = P T    where T, left in the accumulator layout (lane l, reg r = T[4(l/16)+r][l%16]),
//                   is directly the B operand of step r under the same K permutation.
//      The result lands as lane l, reg r = B[Y = 4(l/16)+r][X = l%16].
// G1 (complex, 8 per row) and S (real, 16 per row) occupy the same 64 bytes per (q, ky) row, and
// each row is read and rewritten by the same lane, so they share one LDS buffer.
typedef float floatx4 __attribute__((ext_vector_type(4)));

template <int NP>
struct NgsMfmaShape {
    static constexpr int M = 16;
    static constexpr int SW = 64 / NP;
    static constexpr int RS = 20;             // words per (q, ky) row: 16 used + 4 pad (conflict-free b128 rows)
    static constexpr int QS = M * RS + 16;    // words per subaperture: +16 rotates banks between subapertures
    static constexpr int LDS_F4 = 4 * SW * QS / 4;   // float4 per block (4 waves)
};

// One tile = 4 waves x SW subapertures of frame f.  `lds` holds NgsMfmaShape<NP>::LDS_F4 float4.
// Global loads of one tile (kept separate so that a streaming caller can issue the next
// frame's loads before computing the current one).
template <int NP>
__device__ __forceinline__ void ngs_mfma_load(const NgsArgs& g, int tile, int f, float (&av)[NP], float (&pv)[NP]) {
    using Sh = NgsMfmaShape<NP>;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int q = lane / NP, x = lane - (lane / NP) * NP;
    const int sub = (tile * 4 + w) * Sh::SW + q;
    if (q < Sh::SW && sub < g.n_sub) {
        const int n = g.n_lenses;
        const int i = sub / n, j = sub - (sub / n) * n;
        const int64_t off = (int64_t)(i * NP) * g.ef_size + (int64_t)j * NP + x;
        const float* ar = g.A + f * g.a_stride + off;
        const float* pr = g.phase + f * g.ph_stride + off;
#pragma unroll
        for (int y = 0; y < NP; ++y) {
            av[y] = ar[(int64_t)y * g.ef_size];
            pv[y] = pr[(int64_t)y * g.ef_size];
        }
    } else {
#pragma unroll
        for (int y = 0; y < NP; ++y) { av[y] = 0.f; pv[y] = 0.f; }
    }
}

template <int NP>
__device__ __forceinline__ void ngs_mfma_compute(const NgsArgs& g, int tile, int f, const float (&av)[NP],
                                                 const float (&phv)[NP], float4* lds, double* red) {
    using Sh = NgsMfmaShape<NP>;
    constexpr int M = Sh::M;
    constexpr int SW = Sh::SW;
    constexpr int RS = Sh::RS;
    constexpr int QS = Sh::QS;
    static_assert(NP <= 8, "M = 16 path needs np_sub <= 8");
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    float2* g1 = reinterpret_cast<float2*>(lds + w * (SW * QS / 4));
    float* sS = reinterpret_cast<float*>(lds + w * (SW * QS / 4));
    const int sub0 = (tile * 4 + w) * SW;
    const int n = g.n_lenses;

    {   // phase A: column transforms
        const int q = lane / NP, x = lane - (lane / NP) * NP;
        float2 e[NP];
        const double ax = g.a0 + x;
        const double k_rev = g.k * (1.0 / kTwoPi), tilt_rev = g.tilt * (1.0 / kTwoPi);
#pragma unroll
        for (int y = 0; y < NP; ++y) e[y] = field_sample_rev(av[y], phv[y], k_rev, tilt_rev, ax + g.a0 + y);
        float2 G[M];
        fft_pruned<M, NP, false>(e, G);
        if (q < SW) {
#pragma unroll
            for (int ky = 0; ky < M; ++ky) g1[(q * QS + ky * RS) / 2 + x] = G[ky];
        }
    }
    __syncthreads();

    for (int r = lane; r < SW * M; r += 64) {   // phase B: row transforms and power spectrum
        const int rowoff = (r / M) * QS + (r % M) * RS;
        float2 row[NP];
#pragma unroll
        for (int x = 0; x < NP; ++x) row[x] = g1[rowoff / 2 + x];
        float2 G[M];
        fft_pruned<M, NP, false>(row, G);
        float4* srow = reinterpret_cast<float4*>(sS + rowoff);
#pragma unroll
        for (int c = 0; c < M / 4; ++c)
            srow[c] = make_float4(G[4 * c].x * G[4 * c].x + G[4 * c].y * G[4 * c].y,
                                  G[4 * c + 1].x * G[4 * c + 1].x + G[4 * c + 1].y * G[4 * c + 1].y,
                                  G[4 * c + 2].x * G[4 * c + 2].x + G[4 * c + 2].y * G[4 * c + 2].y,
                                  G[4 * c + 3].x * G[4 * c + 3].x + G[4 * c + 3].y * G[4 * c + 3].y);
    }
    __syncthreads();

    // phase C: binning on the matrix cores
    const int i16 = lane & 15, g4 = lane >> 4;
    const floatx4 qv = *reinterpret_cast<const floatx4*>(g.colw + i16 * M + 4 * g4);
    const floatx4 pv = *reinterpret_cast<const floatx4*>(g.roww + i16 * M + 4 * g4);
    double scale = 1.0;
    if (g.parseval) {
        const double* am = g.amp + (int64_t)f * g.amp_fstep;
        scale = g.S0 * g.pixel_area * am[0] / (g.norm_const * am[1]);
    }
    const int W = n * g.npx;
    double bsum = 0.0;
    for (int q = 0; q < SW; ++q) {
        const int sub = sub0 + q;
        if (sub >= g.n_sub) break;                       // wave-uniform
        const floatx4 sv = *reinterpret_cast<const floatx4*>(sS + q * QS + i16 * RS + 4 * g4);
        floatx4 t = {0.f, 0.f, 0.f, 0.f};
        t = __builtin_amdgcn_mfma_f32_16x16x4f32(sv[0], qv[0], t, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x4f32(sv[1], qv[1], t, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x4f32(sv[2], qv[2], t, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x4f32(sv[3], qv[3], t, 0, 0, 0);
        floatx4 b = {0.f, 0.f, 0.f, 0.f};
        b = __builtin_amdgcn_mfma_f32_16x16x4f32(pv[0], t[0], b, 0, 0, 0);
        b = __builtin_amdgcn_mfma_f32_16x16x4f32(pv[1], t[1], b, 0, 0, 0);
        b = __builtin_amdgcn_mfma_f32_16x16x4f32(pv[2], t[2], b, 0, 0, 0);
        b = __builtin_amdgcn_mfma_f32_16x16x4f32(pv[3], t[3], b, 0, 0, 0);
        // stage the 16 x 16 tile over this subaperture's (already consumed) spectrum in LDS;
        // rows/columns >= npx are exactly zero (zero-padded operators)
        float* bt = sS + q * QS;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            bt[(4 * g4 + r) * RS + i16] = b[r];
            if (!g.parseval && 4 * g4 + r < g.npx && i16 < g.npx) bsum += (double)b[r];
        }
    }
    __syncthreads();
    // coalesced stores: consecutive lanes walk X, then subapertures, then CCD rows
    const int nq = min(SW, g.n_sub - sub0);
    float* ccdf = g.ccd + f * g.ccd_stride;
    if (64 % (4 * SW) == 0 && g.npx == 16 && nq == SW) {
        // lane -> fixed (subaperture q, float4 column x4); CCD rows Y = lane/32 + 2t
        const float fscale = (float)scale;
        const int x4 = lane & 3, q = (lane >> 2) % SW;
        const int sub = sub0 + q;
        const int i = sub / n, j = sub - (sub / n) * n;
        float* base = ccdf + (int64_t)(i * 16) * W + j * 16 + 4 * x4;
        const float* src = sS + q * QS + 4 * x4;
        constexpr int YSTEP = 64 / (4 * SW);
        for (int Y = (lane >> 2) / SW; Y < 16; Y += YSTEP) {
            const float4 v = *reinterpret_cast<const float4*>(src + Y * RS);
            const floatx4 o = {v.x * fscale, v.y * fscale, v.z * fscale, v.w * fscale};
            __builtin_nontemporal_store(o, reinterpret_cast<floatx4*>(base + (int64_t)Y * W));
        }
    } else if (nq > 0) {
        const int per = nq * g.npx * g.npx;
        for (int e = lane; e < per; e += 64) {
            const int X = e % g.npx, t2 = e / g.npx;
            const int q = t2 % nq, Y = t2 / nq;
            const int sub = sub0 + q;
            const int i = sub / n, j = sub - (sub / n) * n;
            ccdf[(int64_t)(i * g.npx + Y) * W + (int64_t)j * g.npx + X] =
                (float)((double)sS[q * QS + Y * RS + X] * scale);
        }
    }
    if (!g.parseval) {
        bsum = block_sum(bsum, red);
        if (threadIdx.x == 0) g.bpart[(int64_t)f * g.n_tiles + tile] = bsum;
    }
}

template <int NP>
__device__ __forceinline__ void ngs_mfma_tile(const NgsArgs& g, int tile, int f, float4* lds, double* red) {
    float av[NP], pv[NP];
    ngs_mfma_load<NP>(g, tile, f, av, pv);
    ngs_mfma_compute<NP>(g, tile, f, av, pv, lds, red);
}

template <int NP>
__global__ void __launch_bounds__(256)
sh_ngs_mfma_kernel(NgsArgs g) {
    __shared__ float4 lds_all[NgsMfmaShape<NP>::LDS_F4];
    __shared__ double red[4];
    ngs_mfma_tile<NP>(g, blockIdx.x, blockIdx.y, lds_all, red);
}

// Streaming form: one block walks FPB consecutive frames of its tile and issues the next frame's
// pupil loads before transforming the current one, so HBM latency overlaps the FFT/MFMA work.
template <int NP, int FPB>
__global__ void __launch_bounds__(256)
sh_ngs_mfma_stream_kernel(NgsArgs g, int n_frames) {
    __shared__ float4 lds_all[NgsMfmaShape<NP>::LDS_F4];
    __shared__ double red[4];
    const int tile = blockIdx.x;
    const int f0 = blockIdx.y * FPB;
    const int f1 = min(n_frames, f0 + FPB);
    float av[NP], pv[NP], an[NP], pn[NP];
    ngs_mfma_load<NP>(g, tile, f0, av, pv);
    for (int f = f0; f < f1; ++f) {
        if (f + 1 < f1) ngs_mfma_load<NP>(g, tile, f + 1, an, pn);
        if (f > f0) __syncthreads();                     // LDS of the previous frame is free
        ngs_mfma_compute<NP>(g, tile, f, av, pv, lds_all, red);
#pragma unroll
        for (int y = 0; y < NP; ++y) { av[y] = an[y]; pv[y] = pn[y]; }
    }
}

// ---------------------------------------------------------------- LGS forward (sh.py:381-394)
// psf = Re IFFT_N(FFT_N(|FFT_N(E)|^2) * K).  With D[m] = (N^2/M^2) FFT_M(S)[m mod M] K[m] on the
// lag grid |m| <= np-1 (S = |FFT_M(E)|^2, M >= 2np-1), the binned image is B = Re(P D Q^T) where
// P, Q (npx x L, complex) already carry N^2/M^2, the rebin, the cut and the mask.
// One block = TS subapertures, 256 threads:
//   1: task (s, x)   column FFT over y of the field                         -> buf[s][ky][x]
//   2: task (s, ky)  pruned FFT over x, S = |.|^2, forward FFT over kx of S  -> buf[s][lx][ky] (L lags)
//   3: task (s, lx)  forward FFT over ky, D = F * K_s, V[Y] = sum_my P[Y][my] D[my]
//   4: task (s, Y, half)  B[Y][X] = Re sum_lx V[Y][lx] Q[X][lx]
template <int NP, int M, int TS>
struct LgsShape {
    static constexpr int L = 2 * NP - 1;
    static constexpr int E1 = M * NP, E2 = L * M, E3 = 16 * L;
    static constexpr int LDS_F2 = TS * (E1 > E2 ? (E1 > E3 ? E1 : E3) : (E2 > E3 ? E2 : E3));
};

template <int NP, int M, int TS>
__device__ __forceinline__ void lgs_tile(const NgsArgs& g, int tile, int f, float2* buf, double* red) {
    constexpr int L = 2 * NP - 1;
    static_assert(L <= M, "lag grid must fit the spectrum grid");
    const int t = threadIdx.x;
    const int n = g.n_lenses;
    const int sub0 = tile * TS;
    const float2* Pc = reinterpret_cast<const float2*>(g.roww);   // npx_pad x L complex
    const float2* Qc = reinterpret_cast<const float2*>(g.colw);

    // phase 1
    for (int task = t; task < TS * NP; task += 256) {
        const int s = task / NP, x = task - (task / NP) * NP;
        const int sub = sub0 + s;
        float2 e[NP];
        if (sub < g.n_sub) {
            const int i = sub / n, j = sub - (sub / n) * n;
            const int64_t off = (int64_t)(i * NP) * g.ef_size + (int64_t)j * NP + x;
            const float* ar = g.A + f * g.a_stride + off;
            const float* pr = g.phase + f * g.ph_stride + off;
            const double k_rev = g.k * (1.0 / kTwoPi), tilt_rev = g.tilt * (1.0 / kTwoPi);
            const double ax = g.a0 + x;
#pragma unroll
            for (int y = 0; y < NP; ++y)
                e[y] = field_sample_rev(ar[(int64_t)y * g.ef_size], pr[(int64_t)y * g.ef_size], k_rev, tilt_rev,
                                        ax + g.a0 + y);
        } else {
#pragma unroll
            for (int y = 0; y < NP; ++y) e[y] = make_float2(0.f, 0.f);
        }
        float2 G[M];
        fft_pruned<M, NP, false>(e, G);
#pragma unroll
        for (int ky = 0; ky < M; ++ky) buf[(s * M + ky) * NP + x] = G[ky];
    }
    __syncthreads();

    // phase 2 (read rows into registers, then overwrite the buffer with lag rows)
    constexpr int T2 = (TS * M + 255) / 256;
    float2 R[T2][L];
#pragma unroll
    for (int h = 0; h < T2; ++h) {
        const int task = t + 256 * h;
        if (task < TS * M) {
            const int s = task / M, ky = task - (task / M) * M;
            float2 row[NP];
#pragma unroll
            for (int x = 0; x < NP; ++x) row[x] = buf[(s * M + ky) * NP + x];
            float2 G[M];
            fft_pruned<M, NP, false>(row, G);
#pragma unroll
            for (int kx = 0; kx < M; ++kx) G[kx] = make_float2(G[kx].x * G[kx].x + G[kx].y * G[kx].y, 0.f);
            fft_radix2<M, false>(G);
#pragma unroll
            for (int li = 0; li < L; ++li) R[h][li] = G[(li - (NP - 1) + M) % M];
        }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < T2; ++h) {
        const int task = t + 256 * h;
        if (task < TS * M) {
            const int s = task / M, ky = task - (task / M) * M;
#pragma unroll
            for (int li = 0; li < L; ++li) buf[(s * L + li) * M + ky] = R[h][li];
        }
    }
    __syncthreads();

    // phase 3: column transforms over ky, kernel multiply, left binning operator
    constexpr int T3 = (TS * L + 255) / 256;
    float2 V[T3][16];
#pragma unroll
    for (int h = 0; h < T3; ++h) {
        const int task = t + 256 * h;
#pragma unroll
        for (int Y = 0; Y < 16; ++Y) V[h][Y] = make_float2(0.f, 0.f);
        if (task < TS * L) {
            const int s = task / L, li = task - (task / L) * L;
            const int sub = sub0 + s;
            float2 F[M];
#pragma unroll
            for (int ky = 0; ky < M; ++ky) F[ky] = buf[(s * L + li) * M + ky];
            fft_radix2<M, false>(F);
            if (sub < g.n_sub) {
                const float2* kr = g.kspec + (int64_t)sub * L * L + li;
#pragma unroll
                for (int mi = 0; mi < L; ++mi) {
                    const float2 d = cmul(F[(mi - (NP - 1) + M) % M], kr[mi * L]);
#pragma unroll
                    for (int Y = 0; Y < 16; ++Y) {          // rows >= npx of P are zero
                        const float2 pc = Pc[Y * L + mi];
                        V[h][Y].x = fmaf(pc.x, d.x, fmaf(-pc.y, d.y, V[h][Y].x));
                        V[h][Y].y = fmaf(pc.x, d.y, fmaf(pc.y, d.x, V[h][Y].y));
                    }
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < T3; ++h) {
        const int task = t + 256 * h;
        if (task < TS * L) {
            const int s = task / L, li = task - (task / L) * L;
#pragma unroll
            for (int Y = 0; Y < 16; ++Y) buf[(s * 16 + Y) * L + li] = V[h][Y];
        }
    }
    __syncthreads();

    // phase 4: right binning operator, real part, scale, store
    double scale = 1.0;
    if (g.parseval) {
        const double* am = g.amp + (int64_t)f * g.amp_fstep;
        scale = g.S0 * g.pixel_area * am[0] / (g.norm_const * am[1]);
    }
    const int W = n * g.npx;
    const int halfw = (g.npx + 1) / 2;
    double bsum = 0.0;
    for (int task = t; task < TS * g.npx * 2; task += 256) {
        const int s = task / (2 * g.npx), rem = task - s * 2 * g.npx;
        const int Y = rem >> 1, hx = rem & 1;
        const int sub = sub0 + s;
        if (sub >= g.n_sub) continue;
        const int i = sub / n, j = sub - (sub / n) * n;
        const float2* vrow = buf + (s * 16 + Y) * L;
        float2 v[L];
#pragma unroll
        for (int li = 0; li < L; ++li) v[li] = vrow[li];
        float* dst = g.ccd + f * g.ccd_stride + (int64_t)(i * g.npx + Y) * W + (int64_t)j * g.npx;
        const int x0 = hx * halfw, x1 = min(g.npx, x0 + halfw);
        for (int X = x0; X < x1; ++X) {
            float acc = 0.f;
#pragma unroll
            for (int li = 0; li < L; ++li) {
                const float2 qc = Qc[X * L + li];
                acc = fmaf(v[li].x, qc.x, fmaf(-v[li].y, qc.y, acc));
            }
            bsum += (double)acc;
            dst[X] = (float)((double)acc * scale);
        }
    }
    if (!g.parseval) {
        bsum = block_sum(bsum, red);
        if (t == 0) g.bpart[(int64_t)f * g.n_tiles + tile] = bsum;
    }
}

template <int NP, int M, int TS>
__global__ void __launch_bounds__(256)
sh_lgs_forward_kernel(NgsArgs g) {
    extern __shared__ float4 lds_raw[];
    __shared__ double red[4];
    lgs_tile<NP, M, TS>(g, blockIdx.x, blockIdx.y, reinterpret_cast<float2*>(lds_raw), red);
}

// Non-Parseval normalisation: scale = S0 * pixel_area * sum(A) / sum(B), then rescale in place.
__global__ void sh_rescale_kernel(float* __restrict__ ccd, int64_t ccd_stride, int64_t n_px,
                                  const double* __restrict__ bpart, int n_tiles,
                                  const double* __restrict__ amp, int amp_fstep,
                                  double S0, double pixel_area) {
    __shared__ double red[16];
    const int f = blockIdx.y;
    double t = 0.0;
    for (int b = threadIdx.x; b < n_tiles; b += blockDim.x) t += bpart[(int64_t)f * n_tiles + b];
    t = block_sum(t, red);
    const double scale = S0 * pixel_area * amp[(int64_t)f * amp_fstep] / t;
    float* c = ccd + f * ccd_stride;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n_px;
         q += (int64_t)gridDim.x * blockDim.x)
        c[q] = (float)((double)c[q] * scale);
}

// ---------------------------------------------------------------- ShSlopec CoG
// One wavefront per subaperture: FP64 moments (flux, thresholded total, x and y numerators).
__global__ void cog_moments_kernel(const float* __restrict__ pix, int64_t pix_stride,
                                   const int32_t* __restrict__ origin, const int32_t* __restrict__ idx,
                                   int n_sub, int np_sub, int frame_w,
                                   const double* __restrict__ xw, const double* __restrict__ yw,
                                   const double* __restrict__ mw, float thr, int thr_mode,
                                   double* __restrict__ mom) {
    const int lane = threadIdx.x % kWave;
    const int sub = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int f = blockIdx.y;
    if (sub >= n_sub) return;
    const float* p = pix + f * pix_stride;
    const int npp = np_sub * np_sub;
    double flux = 0.0, tot = 0.0, sx = 0.0, sy = 0.0;
    for (int q = lane; q < npp; q += kWave) {
        int64_t addr;
        if (origin) {
            const int r = q / np_sub, c = q - (q / np_sub) * np_sub;
            addr = (int64_t)origin[sub] + (int64_t)r * frame_w + c;
        } else {
            addr = idx[(int64_t)sub * npp + q];
        }
        const float v = p[addr];
        flux += (double)v;
        float w;
        if (thr_mode == 1) {              // pedestal: zero below threshold (sh_slopec.py:382-383)
            w = v < thr ? 0.f : v;
        } else {                          // subtract then clamp at zero (sh_slopec.py:385-386)
            w = v - thr;
            w = w < 0.f ? 0.f : w;
        }
        const double wd = (double)w;
        tot += wd * mw[q];
        sx += wd * xw[q];
        sy += wd * yw[q];
    }
    flux = wave_sum(flux);
    tot = wave_sum(tot);
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    if (lane == 0) {
        double* m = mom + ((int64_t)f * n_sub + sub) * 4;
        m[0] = flux; m[1] = tot; m[2] = sx; m[3] = sy;
    }
}

// Fast path for square subaperture blocks with separable weights (xw[r][c] = xv[c],
// yw[r][c] = yv[r], mask_weighted = 1, the ShSlopec default): one lane per pixel row, NP lanes per
// subaperture, subapertures taken in row-major order of their blocks so that the lanes of a wave
// read neighbouring 64-byte row segments.  FP64 accumulation throughout; x weights are
// wave-uniform (scalar loads), y weights multiply the row total.
template <int NP, bool VEC4>
__global__ void __launch_bounds__(256)
cog_rows_kernel(const float* __restrict__ pix, int64_t pix_stride, const int32_t* __restrict__ origin,
                const int32_t* __restrict__ order, int n_sub, int frame_w,
                const double* __restrict__ xv, const double* __restrict__ yv, float thr, int thr_mode,
                double* __restrict__ mom) {
    constexpr int SPW = 64 / NP;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int q = lane / NP, r = lane - (lane / NP) * NP;
    const int k = (blockIdx.x * 4 + w) * SPW + q;
    const int f = blockIdx.y;
    const bool active = k < n_sub;
    double flux = 0.0, tot = 0.0, sx = 0.0;
    if (active) {
        const float* row = pix + f * pix_stride + origin[k] + (int64_t)r * frame_w;
        float v[NP];
        if constexpr (VEC4) {
#pragma unroll
            for (int c = 0; c < NP / 4; ++c) {
                const float4 t = reinterpret_cast<const float4*>(row)[c];
                v[4 * c] = t.x; v[4 * c + 1] = t.y; v[4 * c + 2] = t.z; v[4 * c + 3] = t.w;
            }
        } else {
#pragma unroll
            for (int c = 0; c < NP; ++c) v[c] = row[c];
        }
#pragma unroll
        for (int c = 0; c < NP; ++c) {
            flux += (double)v[c];
            float wv;
            if (thr_mode == 1) {
                wv = v[c] < thr ? 0.f : v[c];
            } else {
                wv = v[c] - thr;
                wv = wv < 0.f ? 0.f : wv;
            }
            tot += (double)wv;
            sx += (double)wv * xv[c];
        }
    }
    double sy = active ? yv[r] * tot : 0.0;
#pragma unroll
    for (int off = NP / 2; off > 0; off >>= 1) {
        flux += __shfl_xor(flux, off, kWave);
        tot += __shfl_xor(tot, off, kWave);
        sx += __shfl_xor(sx, off, kWave);
        sy += __shfl_xor(sy, off, kWave);
    }
    if (active && r == 0) {
        double* m = mom + ((int64_t)f * n_sub + order[k]) * 4;
        m[0] = flux; m[1] = tot; m[2] = sx; m[3] = sy;
    }
}

// One block per frame: mean of the thresholded totals, dim-subaperture cut-off
// (factor = 0 where 1/tot > 1/(1e-3 mean), sh_slopec.py:393-400), slopes and flux outputs.
__global__ void cog_finalize_kernel(const double* __restrict__ mom, int n_sub,
                                    float* __restrict__ slopes, int64_t slopes_stride,
                                    double* __restrict__ flux, double* __restrict__ totals,
                                    double* __restrict__ factor_out) {
    __shared__ double red[16];
    const int f = blockIdx.x;
    const double* m = mom + (int64_t)f * n_sub * 4;
    double st = 0.0, sf = 0.0;
    for (int s = threadIdx.x; s < n_sub; s += blockDim.x) {
        sf += m[4 * s];
        st += m[4 * s + 1];
    }
    sf = block_sum(sf, red);
    st = block_sum(st, red);
    const double mean_tot = st / n_sub;
    const double limit = 1.0 / (mean_tot * 1e-3);
    float* sl = slopes + f * slopes_stride;
    for (int s = threadIdx.x; s < n_sub; s += blockDim.x) {
        double factor = 1.0 / m[4 * s + 1];
        if (factor > limit) factor = 0.0;
        factor_out[(int64_t)f * n_sub + s] = factor;
        if (slopes) {
            sl[s] = (float)(m[4 * s + 2] * factor);
            sl[n_sub + s] = (float)(m[4 * s + 3] * factor);
        }
        if (flux) flux[(int64_t)f * n_sub + s] = m[4 * s];
    }
    if (totals && threadIdx.x == 0) {
        totals[2 * f] = sf;
        totals[2 * f + 1] = sf / n_sub;
    }
}


// ============================================================================ adjoint (VJP)
__global__ void fill_zero_kernel(float* __restrict__ p, int64_t n) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        p[q] = 0.f;
}

// CoG backward (sh_slopec.py:374-404 differentiated with the dim-subaperture factor held fixed):
//   dL/dp = factor * (gx (xw - sx mw) + gy (yw - sy mw)) where the thresholded pixel is alive.
// One wavefront per subaperture; pixels outside every subaperture keep the zero fill.
__global__ void cog_vjp_kernel(const float* __restrict__ pix, int64_t pix_stride,
                               const int32_t* __restrict__ origin, const int32_t* __restrict__ idx,
                               int n_sub, int np_sub, int frame_w,
                               const double* __restrict__ xw, const double* __restrict__ yw,
                               const double* __restrict__ mw, float thr, int thr_mode,
                               const double* __restrict__ mom, const double* __restrict__ factor,
                               const float* __restrict__ gsl, int64_t gsl_stride,
                               float* __restrict__ gpix, int64_t gpix_stride) {
    const int lane = threadIdx.x % kWave;
    const int sub = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int f = blockIdx.y;
    if (sub >= n_sub) return;
    const double fac = factor[(int64_t)f * n_sub + sub];
    const double* m = mom + ((int64_t)f * n_sub + sub) * 4;
    const double sx = m[2] * fac, sy = m[3] * fac;
    const double gx = gsl[f * gsl_stride + sub], gy = gsl[f * gsl_stride + n_sub + sub];
    const int npp = np_sub * np_sub;
    for (int q = lane; q < npp; q += kWave) {
        int64_t addr;
        if (origin) {
            const int r = q / np_sub, c = q - (q / np_sub) * np_sub;
            addr = (int64_t)origin[sub] + (int64_t)r * frame_w + c;
        } else {
            addr = idx[(int64_t)sub * npp + q];
        }
        const float v = pix[f * pix_stride + addr];
        const bool alive = (thr_mode == 1) ? !(v < thr) : (v - thr > 0.f);
        const double gp = alive ? fac * (gx * (xw[q] - sx * mw[q]) + gy * (yw[q] - sy * mw[q])) : 0.0;
        gpix[f * gpix_stride + addr] = (float)gp;
    }
}

struct VjpArgs {
    const float* A;
    int64_t a_stride;
    const float* phase;
    int64_t ph_stride;
    const float* gccd;     // dL/d(CCD intensity), per frame
    int64_t gccd_stride;
    float* gphase;         // dL/d(phase [nm]), per frame
    int64_t gph_stride;
    const float* roww;     // P (rows_alloc x M)
    const float* colw;     // Q
    const double* amp;
    int amp_fstep;
    int n_lenses, ef_size, npx, n_sub;
    double k, tilt, a0, S0, pixel_area, norm_const;
};

// SH backward for the NGS path with Parseval normalisation (I = c B, c independent of the
// phase): with G_B = c dL/dI on the subaperture tile,
//   G_S = P^T G_B Q,   X = G_S conj(F),   H = FFT_M(X) on the np x np support,
//   dL/dphi = -2 (2 pi / lambda) Im(H E).
template <int NP, int M, int TS>
__global__ void __launch_bounds__(TS * NP)
sh_ngs_vjp_kernel(VjpArgs g) {
    extern __shared__ float4 lds_raw[];
    constexpr int G1S = M * NP + 1;
    constexpr int NT = TS * NP;
    constexpr int CPT = (TS * M + NT - 1) / NT;
    float2* buf = reinterpret_cast<float2*>(lds_raw);
    float* gb = reinterpret_cast<float*>(buf + TS * G1S);
    const int npx2 = g.npx * g.npx;
    const int t = threadIdx.x, f = blockIdx.y, tile = blockIdx.x, n = g.n_lenses;
    const double* am = g.amp + (int64_t)f * g.amp_fstep;
    const float c = (float)(g.S0 * g.pixel_area * am[0] / (g.norm_const * am[1]));

    for (int e = t; e < TS * npx2; e += NT) {         // G_B tiles
        const int s = e / npx2, rem = e - s * npx2;
        const int Y = rem / g.npx, X = rem - Y * g.npx;
        const int sub = tile * TS + s;
        float v = 0.f;
        if (sub < g.n_sub) {
            const int i = sub / n, j = sub - (sub / n) * n;
            v = g.gccd[f * g.gccd_stride + (int64_t)(i * g.npx + Y) * (n * g.npx) + j * g.npx + X] * c;
        }
        gb[s * npx2 + rem] = v;
    }
    // row pass: field rows (kept for the last step) and their pruned transforms
    const int s_r = t % TS, y_r = t / TS;
    const int sub_r = tile * TS + s_r;
    float2 e[NP];
    int64_t row_off = 0;
    if (sub_r < g.n_sub) {
        const int i = sub_r / n, j = sub_r - (sub_r / n) * n;
        row_off = (int64_t)(i * NP + y_r) * g.ef_size + (int64_t)j * NP;
        const float* ar = g.A + f * g.a_stride + row_off;
        const float* pr = g.phase + f * g.ph_stride + row_off;
        const double k_rev = g.k * (1.0 / kTwoPi), tilt_rev = g.tilt * (1.0 / kTwoPi);
#pragma unroll
        for (int x = 0; x < NP; ++x)
            e[x] = field_sample_rev(ar[x], pr[x], k_rev, tilt_rev, g.a0 + x + g.a0 + y_r);
    } else {
#pragma unroll
        for (int x = 0; x < NP; ++x) e[x] = make_float2(0.f, 0.f);
    }
    {
        float2 r[M];
        fft_pruned<M, NP, false>(e, r);
#pragma unroll
        for (int kx = 0; kx < M; ++kx) buf[s_r * G1S + kx * NP + y_r] = r[kx];
    }
    __syncthreads();
    // column pass: F, G_S, X, transform over ky (first NP outputs kept)
    float2 Z[CPT][NP];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int cidx = t + q * NT;
#pragma unroll
        for (int y = 0; y < NP; ++y) Z[q][y] = make_float2(0.f, 0.f);
        if (cidx < TS * M) {
            const int s = cidx % TS, kx = cidx / TS;
            float2 col[NP];
#pragma unroll
            for (int y = 0; y < NP; ++y) col[y] = buf[s * G1S + kx * NP + y];
            float2 F[M];
            fft_pruned<M, NP, false>(col, F);
            float gs[M];
#pragma unroll
            for (int ky = 0; ky < M; ++ky) gs[ky] = 0.f;
            const float* gbs = gb + s * npx2;
            for (int Y = 0; Y < g.npx; ++Y) {
                float u = 0.f;
                for (int X = 0; X < g.npx; ++X) u = fmaf(gbs[Y * g.npx + X], g.colw[X * M + kx], u);
#pragma unroll
                for (int ky = 0; ky < M; ++ky) gs[ky] = fmaf(g.roww[Y * M + ky], u, gs[ky]);
            }
            float2 X[M];
#pragma unroll
            for (int ky = 0; ky < M; ++ky) X[ky] = make_float2(gs[ky] * F[ky].x, -gs[ky] * F[ky].y);
            fft_radix2<M, false>(X);
#pragma unroll
            for (int y = 0; y < NP; ++y) Z[q][y] = X[y];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int cidx = t + q * NT;
        if (cidx < TS * M) {
            const int s = cidx % TS, kx = cidx / TS;
#pragma unroll
            for (int y = 0; y < NP; ++y) buf[s * G1S + y * M + kx] = Z[q][y];
        }
    }
    __syncthreads();
    // row pass back: transform over kx, phase gradient
    if (sub_r < g.n_sub) {
        float2 H[M];
#pragma unroll
        for (int kx = 0; kx < M; ++kx) H[kx] = buf[s_r * G1S + y_r * M + kx];
        fft_radix2<M, false>(H);
        float* gr = g.gphase + f * g.gph_stride + row_off;
        const float two_k = (float)(2.0 * g.k);
#pragma unroll
        for (int x = 0; x < NP; ++x) {
            const float im = H[x].x * e[x].y + H[x].y * e[x].x;     // Im(H E)
            gr[x] = -two_k * im;
        }
    }
}

// Interaction matrix from the block-diagonal Jacobian: with fov_ovs == 1 each phase pixel feeds
// one subaperture, so two VJPs (cotangent 1 on every x slope, then on every y slope) give
// Jx, Jy (pupil maps, slope per nm) and IM[m][slot] = scale * sum_{p in subap} J[p] mode_m[p]
// (scale 2 reproduces the push-pull convention (s(+a) - s(-a))/a of im_rec_calibrator.py:58-63).
// Block = one strip of NP pupil rows x a chunk of modes; one thread per pupil column.
template <int NP>
__global__ void im_from_jacobian_kernel(const float* __restrict__ jx, const float* __restrict__ jy, int ef,
                                        int n_lenses, const float* __restrict__ modes, int64_t mode_stride,
                                        int n_modes, int modes_per_block, const int32_t* __restrict__ slot,
                                        int n_valid, float scale, float* __restrict__ im) {
    const int i = blockIdx.x;
    const int col = threadIdx.x;
    const bool active = col < ef;
    float jxr[NP], jyr[NP];
#pragma unroll
    for (int r = 0; r < NP; ++r) {
        const int64_t o = (int64_t)(i * NP + r) * ef + col;
        jxr[r] = active ? jx[o] : 0.f;
        jyr[r] = active ? jy[o] : 0.f;
    }
    const int j = col / NP, xin = col - j * NP;
    const int sl = (active && xin == 0) ? slot[i * n_lenses + j] : -1;
    const int m0 = blockIdx.y * modes_per_block;
    const int m1 = min(n_modes, m0 + modes_per_block);
    for (int m = m0; m < m1; ++m) {
        float px = 0.f, py = 0.f;
        if (active) {
            const float* md = modes + m * mode_stride + (int64_t)(i * NP) * ef + col;
#pragma unroll
            for (int r = 0; r < NP; ++r) {
                const float v = md[(int64_t)r * ef];
                px = fmaf(jxr[r], v, px);
                py = fmaf(jyr[r], v, py);
            }
        }
#pragma unroll
        for (int off = NP / 2; off > 0; off >>= 1) {
            px += __shfl_xor(px, off, kWave);
            py += __shfl_xor(py, off, kWave);
        }
        if (sl >= 0) {
            im[(int64_t)m * 2 * n_valid + sl] = scale * px;
            im[(int64_t)m * 2 * n_valid + n_valid + sl] = scale * py;
        }
    }
}


// ---------------------------------------------------------------- multi-sensor persistent batch
// One launch processes a work list of (sensor, frame, tile) items of uneven shapes: NGS tiles
// (np 8, MFMA binning, 32 subapertures) and LGS tiles (np 16, N 128, 8 subapertures).  Blocks
// stride over the list; LDS is sized for the larger body.  Per-sensor completion times are
// recorded with the constant-rate wall clock for latency statistics.
struct BatchItem {
    int32_t sensor, frame, tile, kind;   // kind 0: NGS MFMA np 8, 1: LGS np 16 / M 32
};

constexpr int kBatchLdsBytes =
    (NgsMfmaShape<8>::LDS_F4 * 16 > LgsShape<16, 32, 8>::LDS_F2 * 8) ? NgsMfmaShape<8>::LDS_F4 * 16
                                                                        : LgsShape<16, 32, 8>::LDS_F2 * 8;

__global__ void batch_stamp_kernel(unsigned long long* __restrict__ ts, int n_sensors) {
    if (threadIdx.x == 0) ts[0] = (unsigned long long)wall_clock64();
    for (int s = threadIdx.x; s < n_sensors; s += blockDim.x) ts[1 + s] = 0ull;
}

__global__ void __launch_bounds__(256)
sh_batch_kernel(const NgsArgs* __restrict__ sensors, const BatchItem* __restrict__ items, int n_items,
                unsigned long long* __restrict__ ts) {
    extern __shared__ float4 lds_raw[];
    __shared__ double red[4];
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        __syncthreads();                               // LDS of the previous item is free
        const BatchItem w = items[it];
        const NgsArgs& g = sensors[w.sensor];
        if (w.kind == 0)
            ngs_mfma_tile<8>(g, w.tile, w.frame, lds_raw, red);
        else
            lgs_tile<16, 32, 8>(g, w.tile, w.frame, reinterpret_cast<float2*>(lds_raw), red);
        __syncthreads();
        if (threadIdx.x == 0) atomicMax(&ts[1 + w.sensor], (unsigned long long)wall_clock64());
    }
}

}  // namespace

// ============================================================================ plans
struct shwfs_plan {
    shwfs_geom g{};
    int device = 0;
    int n_sub = 0;
    int npx_pad = 0;
    int ts = 0;
    int n_tiles = 0;
    int amp_blocks = 0;
    int rows_alloc = 0;            // rows of the uploaded binning operators (zero padded)
    bool mfma = false;             // binning on the matrix cores (sh_ngs_mfma_kernel)
    bool kernels_set = false;      // LGS: kernel spectra uploaded
    int fpb = 4;                   // frames per block of the streaming MFMA kernel (1: no streaming)
    int L = 0;                     // LGS lag-grid size 2 np - 1
    float* d_roww = nullptr;
    float* d_colw = nullptr;
    float* d_kspec = nullptr;
    double* d_amp = nullptr;       // [max_frames][2]
    double* d_amp_part = nullptr;  // [max_frames][amp_blocks][2]
    unsigned* d_counter = nullptr; // [max_frames]
    double* d_bpart = nullptr;     // [max_frames][n_tiles]
    hipEvent_t ev_start = nullptr; // optional timing hook around the main kernel
    hipEvent_t ev_stop = nullptr;
};

struct shslopec_plan {
    int device = 0;
    int n_sub = 0, np_sub = 0, frame_h = 0, frame_w = 0, max_frames = 0;
    bool rows_path = false;        // cog_rows_kernel (separable weights, square blocks)
    bool vec4 = false;
    int32_t* d_sorted = nullptr;   // row-major sorted block origins
    int32_t* d_order = nullptr;    // subaperture index of each sorted slot
    double* d_sep = nullptr;       // xv | yv
    int32_t* d_origin = nullptr;
    int32_t* d_idx = nullptr;
    double* d_w = nullptr;     // xw | yw | mw
    double* d_mom = nullptr;   // [max_frames][n_sub][4]
    double* d_factor = nullptr;  // [max_frames][n_sub]: 1/tot or 0 (dim-subaperture cut-off)
};

namespace {

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

template <int NP, int M>
int launch_ngs(shwfs_plan* p, const NgsArgs& a, int n_frames, hipStream_t st) {
    constexpr int TS = 256 / NP;
    const int g1_bytes = TS * (M * NP + 1) * (int)sizeof(float2);
    const int u_bytes = TS * (p->npx_pad * M + 4) * (int)sizeof(float);
    const int lds = (g1_bytes > u_bytes ? g1_bytes : u_bytes);
    if (lds > 150 * 1024) return set_error(SHWFS_ERR_UNSUPPORTED, "LDS budget exceeded for this geometry");
    dim3 grid(p->n_tiles, n_frames);
    hipLaunchKernelGGL((sh_ngs_forward_kernel<NP, M, TS>), grid, dim3(TS * NP), lds, st, a);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

int ngs_tile_size(int np) { return 256 / np; }

template <int NP>
int launch_ngs_mfma(shwfs_plan* p, const NgsArgs& a, int n_frames, hipStream_t st) {
    const int fpb = p->fpb;
    if (fpb <= 1 || n_frames == 1) {
        hipLaunchKernelGGL((sh_ngs_mfma_kernel<NP>), dim3(p->n_tiles, n_frames), dim3(256), 0, st, a);
    } else {
        dim3 grid(p->n_tiles, (n_frames + fpb - 1) / fpb);
        switch (fpb) {
            case 2: hipLaunchKernelGGL((sh_ngs_mfma_stream_kernel<NP, 2>), grid, dim3(256), 0, st, a, n_frames); break;
            case 4: hipLaunchKernelGGL((sh_ngs_mfma_stream_kernel<NP, 4>), grid, dim3(256), 0, st, a, n_frames); break;
            default: hipLaunchKernelGGL((sh_ngs_mfma_stream_kernel<NP, 8>), grid, dim3(256), 0, st, a, n_frames); break;
        }
    }
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

using NgsLauncher = int (*)(shwfs_plan*, const NgsArgs&, int, hipStream_t);

template <int NP, int M, int TS>
int launch_lgs(shwfs_plan* p, const NgsArgs& a, int n_frames, hipStream_t st) {
    constexpr int lds = LgsShape<NP, M, TS>::LDS_F2 * (int)sizeof(float2);
    static_assert(lds <= 64 * 1024, "LGS tile exceeds 64 KiB of LDS");
    dim3 grid(p->n_tiles, n_frames);
    hipLaunchKernelGGL((sh_lgs_forward_kernel<NP, M, TS>), grid, dim3(256), lds, st, a);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

NgsLauncher pick_lgs(int np, int M, int* ts) {
    if (np == 16 && M == 32) { *ts = 8; return launch_lgs<16, 32, 8>; }
    if (np == 12 && M == 32) { *ts = 8; return launch_lgs<12, 32, 8>; }
    if (np == 8 && M == 16) { *ts = 16; return launch_lgs<8, 16, 16>; }
    if (np == 6 && M == 16) { *ts = 16; return launch_lgs<6, 16, 16>; }
    return nullptr;
}

// MFMA binning path: M = 16 (np_sub 5..8) and at most 16 CCD px per subaperture.
NgsLauncher pick_ngs_mfma(int np, int M, int npx) {
    if (M != 16 || npx > 16) return nullptr;
    switch (np) {
        case 5: return launch_ngs_mfma<5>;
        case 6: return launch_ngs_mfma<6>;
        case 7: return launch_ngs_mfma<7>;
        case 8: return launch_ngs_mfma<8>;
        default: return nullptr;
    }
}

NgsLauncher pick_ngs(int np, int M) {
    switch (np) {
        case 4: return M == 8 ? launch_ngs<4, 8> : nullptr;
        case 5: return M == 16 ? launch_ngs<5, 16> : nullptr;
        case 6: return M == 16 ? launch_ngs<6, 16> : nullptr;
        case 8: return M == 16 ? launch_ngs<8, 16> : nullptr;
        case 10: return M == 32 ? launch_ngs<10, 32> : nullptr;
        case 12: return M == 32 ? launch_ngs<12, 32> : nullptr;
        case 16: return M == 32 ? launch_ngs<16, 32> : nullptr;
        default: return nullptr;
    }
}

}  // namespace

extern "C" {

const char* shwfs_version(void) { return "shwfs-gfx950 0.1"; }

const char* shwfs_last_error(void) { return g_last_error.c_str(); }

int shwfs_plan_create(const shwfs_geom* geom, const double* row_w, const double* col_w,
                      int device, shwfs_plan** out) {
    if (!geom || !row_w || !col_w || !out) return set_error(SHWFS_ERR_ARG, "null argument");
    const shwfs_geom& g = *geom;
    if (g.n_lenses <= 0 || g.np_sub <= 0 || g.subap_npx <= 0 || g.max_frames <= 0)
        return set_error(SHWFS_ERR_ARG, "non-positive geometry field");
    if (g.ef_size != g.n_lenses * g.np_sub)
        return set_error(SHWFS_ERR_ARG, "ef_size must equal n_lenses * np_sub");
    if (g.spec_size < 2 * g.np_sub - 1)
        return set_error(SHWFS_ERR_ARG, "spec_size must be >= 2*np_sub-1");
    if (g.variant < 0 || g.variant > 2) return set_error(SHWFS_ERR_ARG, "variant must be 0, 1 or 2");
    NgsLauncher mf = nullptr, v1 = nullptr, lg = nullptr;
    int lgs_ts = 0;
    bool use_mfma = false;
    if (g.lgs) {
        lg = pick_lgs(g.np_sub, g.spec_size, &lgs_ts);
        if (!lg) return set_error(SHWFS_ERR_UNSUPPORTED, "LGS path instantiated for (np_sub, spec_size) in {(16,32),(12,32),(8,16),(6,16)}");
        if (g.subap_npx > 16) return set_error(SHWFS_ERR_UNSUPPORTED, "LGS path needs subap_npx <= 16");
    } else {
        mf = pick_ngs_mfma(g.np_sub, g.spec_size, g.subap_npx);
        v1 = pick_ngs(g.np_sub, g.spec_size);
        use_mfma = (g.variant == 2) || (g.variant == 0 && mf != nullptr);
        if (use_mfma && !mf) return set_error(SHWFS_ERR_UNSUPPORTED, "MFMA variant needs spec_size 16, np_sub 5..8, subap_npx <= 16");
        if (!use_mfma && !v1)
            return set_error(SHWFS_ERR_UNSUPPORTED, "np_sub / spec_size combination not instantiated");
    }
    DeviceGuard dg(device);
    auto* p = new shwfs_plan();
    p->g = g;
    p->device = device;
    p->n_sub = g.n_lenses * g.n_lenses;
    p->npx_pad = (g.subap_npx + 3) / 4 * 4;
    p->mfma = use_mfma;
    p->L = 2 * g.np_sub - 1;
    p->amp_blocks = 64;
    const int M = g.spec_size;
    std::vector<float> rw, cw;
    if (g.lgs) {
        p->rows_alloc = 16;
        p->ts = lgs_ts;
        rw.assign((size_t)16 * p->L * 2, 0.f);
        cw.assign((size_t)16 * p->L * 2, 0.f);
        for (size_t q = 0; q < (size_t)g.subap_npx * p->L * 2; ++q) {
            rw[q] = (float)row_w[q];
            cw[q] = (float)col_w[q];
        }
    } else {
        p->rows_alloc = use_mfma ? 16 : p->npx_pad;
        p->ts = use_mfma ? 4 * (64 / g.np_sub) : ngs_tile_size(g.np_sub);
        rw.assign((size_t)p->rows_alloc * M, 0.f);
        cw.assign((size_t)p->rows_alloc * M, 0.f);
        for (int Y = 0; Y < g.subap_npx; ++Y)
            for (int k = 0; k < M; ++k) {
                rw[(size_t)Y * M + k] = (float)row_w[(size_t)Y * M + k];
                cw[(size_t)Y * M + k] = (float)col_w[(size_t)Y * M + k];
            }
    }
    p->n_tiles = (p->n_sub + p->ts - 1) / p->ts;
    auto cleanup = [&](int code) { shwfs_plan_destroy(p); return code; };
    hipError_t e = hipSuccess;
#define PLAN_ALLOC(ptr, bytes)                                                                  \
    e = hipMalloc((void**)&(ptr), (bytes));                                                     \
    if (e != hipSuccess) return cleanup(set_error(SHWFS_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e)));
    PLAN_ALLOC(p->d_roww, rw.size() * sizeof(float));
    PLAN_ALLOC(p->d_colw, cw.size() * sizeof(float));
    PLAN_ALLOC(p->d_amp, (size_t)g.max_frames * 2 * sizeof(double));
    PLAN_ALLOC(p->d_amp_part, (size_t)g.max_frames * p->amp_blocks * 2 * sizeof(double));
    PLAN_ALLOC(p->d_counter, (size_t)g.max_frames * sizeof(unsigned));
    PLAN_ALLOC(p->d_bpart, (size_t)g.max_frames * p->n_tiles * sizeof(double));
    if (g.lgs) PLAN_ALLOC(p->d_kspec, (size_t)p->n_sub * p->L * p->L * 2 * sizeof(float));
#undef PLAN_ALLOC
    if ((e = hipMemcpy(p->d_roww, rw.data(), rw.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(p->d_colw, cw.data(), cw.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(p->d_counter, 0, (size_t)g.max_frames * sizeof(unsigned))) != hipSuccess)
        return cleanup(set_error(SHWFS_ERR_HIP, std::string("plan upload: ") + hipGetErrorString(e)));
    *out = p;
    return SHWFS_OK;
}

int shwfs_plan_destroy(shwfs_plan* p) {
    if (!p) return SHWFS_OK;
    DeviceGuard dg(p->device);
    (void)hipFree(p->d_roww);
    (void)hipFree(p->d_colw);
    (void)hipFree(p->d_kspec);
    (void)hipFree(p->d_amp);
    (void)hipFree(p->d_amp_part);
    (void)hipFree(p->d_counter);
    (void)hipFree(p->d_bpart);
    delete p;
    return SHWFS_OK;
}

int shwfs_plan_set_frames_per_block(shwfs_plan* p, int fpb) {
    if (!p) return set_error(SHWFS_ERR_ARG, "null plan");
    if (fpb != 1 && fpb != 2 && fpb != 4 && fpb != 8) return set_error(SHWFS_ERR_ARG, "frames per block must be 1, 2, 4 or 8");
    p->fpb = fpb;
    return SHWFS_OK;
}

int shwfs_plan_set_timing_events(shwfs_plan* p, void* start, void* stop) {
    if (!p) return set_error(SHWFS_ERR_ARG, "null plan");
    if ((start == nullptr) != (stop == nullptr)) return set_error(SHWFS_ERR_ARG, "give both events or neither");
    p->ev_start = reinterpret_cast<hipEvent_t>(start);
    p->ev_stop = reinterpret_cast<hipEvent_t>(stop);
    return SHWFS_OK;
}

int shwfs_set_kernels(shwfs_plan* p, const float* kspec, void* stream) {
    if (!p || !kspec) return set_error(SHWFS_ERR_ARG, "null argument");
    if (!p->g.lgs) return set_error(SHWFS_ERR_ARG, "plan was not created for the LGS path");
    DeviceGuard dg(p->device);
    SHWFS_HIP_CHECK(hipMemcpyAsync(p->d_kspec, kspec, (size_t)p->n_sub * p->L * p->L * 2 * sizeof(float),
                                   hipMemcpyDeviceToDevice, reinterpret_cast<hipStream_t>(stream)));
    p->kernels_set = true;
    return SHWFS_OK;
}

int shwfs_forward(shwfs_plan* p, int n_frames, const float* A, int64_t a_stride,
                  const float* phase, int64_t ph_stride, double S0, float* ccd,
                  int64_t ccd_stride, void* stream) {
    if (!p || !A || !phase || !ccd) return set_error(SHWFS_ERR_ARG, "null argument");
    if (n_frames <= 0 || n_frames > p->g.max_frames)
        return set_error(SHWFS_ERR_ARG, "n_frames outside [1, max_frames]");
    const shwfs_geom& g = p->g;
    const int64_t ef_px = (int64_t)g.ef_size * g.ef_size;
    const int64_t ccd_side = (int64_t)g.n_lenses * g.subap_npx;
    if (a_stride != 0 && a_stride < ef_px) return set_error(SHWFS_ERR_ARG, "A frame stride too small");
    if (n_frames > 1 && ph_stride < ef_px) return set_error(SHWFS_ERR_ARG, "phase frame stride too small");
    if (n_frames > 1 && ccd_stride < ccd_side * ccd_side)
        return set_error(SHWFS_ERR_ARG, "ccd frame stride too small");
    DeviceGuard dg(p->device);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);

    // amplitude statistics: one pass per distinct amplitude frame
    const int amp_frames = (a_stride == 0) ? 1 : n_frames;
    hipLaunchKernelGGL(amp_stats_kernel, dim3(p->amp_blocks, amp_frames), dim3(256), 0, st,
                       A, a_stride, ef_px, p->d_amp_part, p->d_counter, p->d_amp);
    SHWFS_LAUNCH_CHECK();

    NgsArgs a{};
    a.A = A; a.a_stride = a_stride;
    a.phase = phase; a.ph_stride = ph_stride;
    a.ccd = ccd; a.ccd_stride = ccd_stride;
    a.roww = p->d_roww; a.colw = p->d_colw;
    a.amp = p->d_amp; a.amp_fstep = (a_stride == 0) ? 0 : 2;
    a.bpart = p->d_bpart;
    a.n_lenses = g.n_lenses; a.ef_size = g.ef_size; a.npx = g.subap_npx; a.npx_pad = p->npx_pad;
    a.n_sub = p->n_sub; a.n_tiles = p->n_tiles; a.parseval = g.parseval;
    a.k = g.k_rad_per_nm; a.tilt = g.tilt_rad; a.a0 = g.tilt_a0;
    a.S0 = S0; a.pixel_area = g.pixel_area; a.norm_const = g.norm_const;
    a.kspec = reinterpret_cast<const float2*>(p->d_kspec);
    if (g.lgs && !p->kernels_set) return set_error(SHWFS_ERR_ARG, "LGS plan: call shwfs_set_kernels first");
    int lts = 0;
    NgsLauncher L = g.lgs ? pick_lgs(g.np_sub, g.spec_size, &lts)
                          : (p->mfma ? pick_ngs_mfma(g.np_sub, g.spec_size, g.subap_npx) : pick_ngs(g.np_sub, g.spec_size));
    if (p->ev_start) SHWFS_HIP_CHECK(hipEventRecord(p->ev_start, st));
    int rc = L(p, a, n_frames, st);
    if (rc) return rc;
    if (p->ev_stop) SHWFS_HIP_CHECK(hipEventRecord(p->ev_stop, st));
    if (!g.parseval) {
        hipLaunchKernelGGL(sh_rescale_kernel, dim3(64, n_frames), dim3(256), 0, st, ccd, ccd_stride,
                           ccd_side * ccd_side, p->d_bpart, p->n_tiles, p->d_amp,
                           (a_stride == 0) ? 0 : 2, S0, g.pixel_area);
        SHWFS_LAUNCH_CHECK();
    }
    return SHWFS_OK;
}

int shslopec_plan_create(int n_sub, int np_sub, int frame_h, int frame_w,
                         const int32_t* block_origin, const int32_t* idx,
                         const double* xw, const double* yw, const double* mw,
                         int max_frames, int device, shslopec_plan** out) {
    if (!out || !xw || !yw || !mw) return set_error(SHWFS_ERR_ARG, "null argument");
    if (n_sub <= 0 || np_sub <= 0 || max_frames <= 0 || frame_h <= 0 || frame_w <= 0)
        return set_error(SHWFS_ERR_ARG, "non-positive size");
    if (!block_origin && !idx) return set_error(SHWFS_ERR_ARG, "need block_origin or idx");
    const int64_t npp = (int64_t)np_sub * np_sub;
    const int64_t frame_px = (int64_t)frame_h * frame_w;
    // host-side bounds check of every pixel address the kernel will read
    if (block_origin) {
        for (int s = 0; s < n_sub; ++s) {
            const int64_t o = block_origin[s];
            const int64_t r0 = o / frame_w, c0 = o % frame_w;
            if (o < 0 || r0 + np_sub > frame_h || c0 + np_sub > frame_w)
                return set_error(SHWFS_ERR_ARG, "subaperture block outside the pixel frame");
        }
    } else {
        for (int64_t q = 0; q < (int64_t)n_sub * npp; ++q)
            if (idx[q] < 0 || idx[q] >= frame_px) return set_error(SHWFS_ERR_ARG, "pixel index out of range");
    }
    DeviceGuard dg(device);
    auto* p = new shslopec_plan();
    p->device = device; p->n_sub = n_sub; p->np_sub = np_sub;
    p->frame_h = frame_h; p->frame_w = frame_w; p->max_frames = max_frames;
    auto cleanup = [&](int code) { shslopec_plan_destroy(p); return code; };
    hipError_t e;
    std::vector<double> w(3 * npp);
    memcpy(w.data(), xw, npp * sizeof(double));
    memcpy(w.data() + npp, yw, npp * sizeof(double));
    memcpy(w.data() + 2 * npp, mw, npp * sizeof(double));
    if ((e = hipMalloc((void**)&p->d_w, w.size() * sizeof(double))) != hipSuccess ||
        (e = hipMalloc((void**)&p->d_mom, (size_t)max_frames * n_sub * 4 * sizeof(double))) != hipSuccess ||
        (e = hipMalloc((void**)&p->d_factor, (size_t)max_frames * n_sub * sizeof(double))) != hipSuccess ||
        (e = hipMemcpy(p->d_w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(set_error(SHWFS_ERR_HIP, std::string("slopec plan: ") + hipGetErrorString(e)));
    if (block_origin) {
        if ((e = hipMalloc((void**)&p->d_origin, n_sub * sizeof(int32_t))) != hipSuccess ||
            (e = hipMemcpy(p->d_origin, block_origin, n_sub * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess)
            return cleanup(set_error(SHWFS_ERR_HIP, std::string("slopec plan: ") + hipGetErrorString(e)));
        // separable default weights -> row kernel with row-major processing order
        bool sep = (np_sub == 4 || np_sub == 8 || np_sub == 16);
        for (int64_t r = 0; sep && r < np_sub; ++r)
            for (int64_t c = 0; c < np_sub; ++c) {
                const int64_t q = r * np_sub + c;
                if (mw[q] != 1.0 || xw[q] != xw[c] || yw[q] != yw[r * np_sub]) { sep = false; break; }
            }
        if (sep) {
            std::vector<int32_t> order(n_sub), sorted(n_sub);
            for (int s = 0; s < n_sub; ++s) order[s] = s;
            std::stable_sort(order.begin(), order.end(),
                             [&](int32_t a, int32_t b) { return block_origin[a] < block_origin[b]; });
            bool aligned = (frame_w % 4 == 0) && (np_sub % 4 == 0);
            for (int s = 0; s < n_sub; ++s) {
                sorted[s] = block_origin[order[s]];
                if (sorted[s] % 4 != 0) aligned = false;
            }
            std::vector<double> sepw(2 * np_sub);
            for (int c = 0; c < np_sub; ++c) sepw[c] = xw[c];
            for (int r = 0; r < np_sub; ++r) sepw[np_sub + r] = yw[(int64_t)r * np_sub];
            if ((e = hipMalloc((void**)&p->d_sorted, n_sub * sizeof(int32_t))) != hipSuccess ||
                (e = hipMalloc((void**)&p->d_order, n_sub * sizeof(int32_t))) != hipSuccess ||
                (e = hipMalloc((void**)&p->d_sep, sepw.size() * sizeof(double))) != hipSuccess ||
                (e = hipMemcpy(p->d_sorted, sorted.data(), n_sub * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(p->d_order, order.data(), n_sub * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(p->d_sep, sepw.data(), sepw.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
                return cleanup(set_error(SHWFS_ERR_HIP, std::string("slopec plan: ") + hipGetErrorString(e)));
            p->rows_path = true;
            p->vec4 = aligned;
        }
    } else {
        if ((e = hipMalloc((void**)&p->d_idx, n_sub * npp * sizeof(int32_t))) != hipSuccess ||
            (e = hipMemcpy(p->d_idx, idx, n_sub * npp * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess)
            return cleanup(set_error(SHWFS_ERR_HIP, std::string("slopec plan: ") + hipGetErrorString(e)));
    }
    *out = p;
    return SHWFS_OK;
}

int shslopec_plan_destroy(shslopec_plan* p) {
    if (!p) return SHWFS_OK;
    DeviceGuard dg(p->device);
    (void)hipFree(p->d_origin);
    (void)hipFree(p->d_idx);
    (void)hipFree(p->d_w);
    (void)hipFree(p->d_mom);
    (void)hipFree(p->d_factor);
    (void)hipFree(p->d_sorted);
    (void)hipFree(p->d_order);
    (void)hipFree(p->d_sep);
    delete p;
    return SHWFS_OK;
}

int shslopec_forward(shslopec_plan* p, int n_frames, const float* pixels, int64_t pix_stride,
                     float thr, int thr_mode, float* slopes, int64_t slopes_stride,
                     double* flux, double* totals, void* stream) {
    if (!p || !pixels || !slopes || !flux || !totals) return set_error(SHWFS_ERR_ARG, "null argument");
    if (n_frames <= 0 || n_frames > p->max_frames)
        return set_error(SHWFS_ERR_ARG, "n_frames outside [1, max_frames]");
    if (thr_mode != 0 && thr_mode != 1) return set_error(SHWFS_ERR_ARG, "thr_mode must be 0 or 1");
    if (n_frames > 1 && pix_stride < (int64_t)p->frame_h * p->frame_w)
        return set_error(SHWFS_ERR_ARG, "pixel frame stride too small");
    if (n_frames > 1 && slopes_stride < 2 * (int64_t)p->n_sub)
        return set_error(SHWFS_ERR_ARG, "slopes frame stride too small");
    DeviceGuard dg(p->device);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int npp = p->np_sub * p->np_sub;
    const double* xw = p->d_w;
    const double* yw = p->d_w + npp;
    const double* mw = p->d_w + 2 * npp;
    if (p->rows_path) {
        const int spb = 4 * (64 / p->np_sub);
        dim3 grid((p->n_sub + spb - 1) / spb, n_frames);
        const double* xv = p->d_sep;
        const double* yv = p->d_sep + p->np_sub;
#define COG_ROWS(NPV, V4) hipLaunchKernelGGL((cog_rows_kernel<NPV, V4>), grid, dim3(256), 0, st, pixels, pix_stride, \
                                             p->d_sorted, p->d_order, p->n_sub, p->frame_w, xv, yv, thr, thr_mode, p->d_mom)
        switch (p->np_sub) {
            case 4: if (p->vec4) COG_ROWS(4, true); else COG_ROWS(4, false); break;
            case 8: if (p->vec4) COG_ROWS(8, true); else COG_ROWS(8, false); break;
            default: if (p->vec4) COG_ROWS(16, true); else COG_ROWS(16, false); break;
        }
#undef COG_ROWS
    } else {
        constexpr int kSubPerBlock = 4;
        dim3 grid((p->n_sub + kSubPerBlock - 1) / kSubPerBlock, n_frames);
        hipLaunchKernelGGL(cog_moments_kernel, grid, dim3(kSubPerBlock * kWave), 0, st, pixels, pix_stride,
                           p->d_origin, p->d_idx, p->n_sub, p->np_sub, p->frame_w, xw, yw, mw, thr,
                           thr_mode, p->d_mom);
    }
    SHWFS_LAUNCH_CHECK();
    hipLaunchKernelGGL(cog_finalize_kernel, dim3(n_frames), dim3(1024), 0, st, p->d_mom, p->n_sub,
                       slopes, slopes_stride, flux, totals, p->d_factor);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

}  // extern "C"

namespace {
template <int NP, int M, int TS>
int launch_vjp(const VjpArgs& a, int n_tiles, int n_frames, int npx, hipStream_t st) {
    const int lds = TS * (M * NP + 1) * (int)sizeof(float2) + TS * npx * npx * (int)sizeof(float);
    if (lds > 150 * 1024) return set_error(SHWFS_ERR_UNSUPPORTED, "VJP tile exceeds the LDS budget");
    hipLaunchKernelGGL((sh_ngs_vjp_kernel<NP, M, TS>), dim3(n_tiles, n_frames), dim3(TS * NP), lds, st, a);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}
using VjpLauncher = int (*)(const VjpArgs&, int, int, int, hipStream_t);
VjpLauncher pick_vjp(int np, int M, int* ts) {
    if (np == 8 && M == 16) { *ts = 32; return launch_vjp<8, 16, 32>; }
    if (np == 6 && M == 16) { *ts = 42; return launch_vjp<6, 16, 42>; }
    if (np == 4 && M == 8) { *ts = 64; return launch_vjp<4, 8, 64>; }
    if (np == 16 && M == 32) { *ts = 16; return launch_vjp<16, 32, 16>; }
    return nullptr;
}
}  // namespace

extern "C" {

int shwfs_vjp(shwfs_plan* p, int n_frames, const float* A, int64_t a_stride, const float* phase,
              int64_t ph_stride, double S0, const float* grad_ccd, int64_t gccd_stride, float* grad_phase,
              int64_t gph_stride, void* stream) {
    if (!p || !A || !phase || !grad_ccd || !grad_phase) return set_error(SHWFS_ERR_ARG, "null argument");
    const shwfs_geom& g = p->g;
    if (g.lgs) return set_error(SHWFS_ERR_UNSUPPORTED, "VJP is available for the NGS path");
    if (!g.parseval)
        return set_error(SHWFS_ERR_UNSUPPORTED, "VJP needs the Parseval normalisation (no FoV cut, all-ones mask)");
    if (n_frames <= 0 || n_frames > g.max_frames) return set_error(SHWFS_ERR_ARG, "n_frames outside [1, max_frames]");
    const int64_t ef_px = (int64_t)g.ef_size * g.ef_size;
    const int64_t ccd_px = (int64_t)g.n_lenses * g.subap_npx * g.n_lenses * g.subap_npx;
    if (n_frames > 1 && ((a_stride != 0 && a_stride < ef_px) || (ph_stride != 0 && ph_stride < ef_px) ||
                         gccd_stride < ccd_px || gph_stride < ef_px))
        return set_error(SHWFS_ERR_ARG, "frame stride too small");
    int ts = 0;
    VjpLauncher L = pick_vjp(g.np_sub, g.spec_size, &ts);
    if (!L) return set_error(SHWFS_ERR_UNSUPPORTED, "VJP instantiated for (np_sub, spec_size) in {(4,8),(6,16),(8,16),(16,32)}");
    DeviceGuard dg(p->device);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int amp_frames = (a_stride == 0) ? 1 : n_frames;
    hipLaunchKernelGGL(amp_stats_kernel, dim3(p->amp_blocks, amp_frames), dim3(256), 0, st,
                       A, a_stride, ef_px, p->d_amp_part, p->d_counter, p->d_amp);
    SHWFS_LAUNCH_CHECK();
    VjpArgs a{};
    a.A = A; a.a_stride = a_stride; a.phase = phase; a.ph_stride = ph_stride;
    a.gccd = grad_ccd; a.gccd_stride = gccd_stride; a.gphase = grad_phase; a.gph_stride = gph_stride;
    a.roww = p->d_roww; a.colw = p->d_colw; a.amp = p->d_amp; a.amp_fstep = (a_stride == 0) ? 0 : 2;
    a.n_lenses = g.n_lenses; a.ef_size = g.ef_size; a.npx = g.subap_npx; a.n_sub = p->n_sub;
    a.k = g.k_rad_per_nm; a.tilt = g.tilt_rad; a.a0 = g.tilt_a0; a.S0 = S0; a.pixel_area = g.pixel_area;
    a.norm_const = g.norm_const;
    const int n_tiles = (p->n_sub + ts - 1) / ts;
    return L(a, n_tiles, n_frames, g.subap_npx, st);
}

int shslopec_vjp(shslopec_plan* p, int n_frames, const float* pixels, int64_t pix_stride, float thr, int thr_mode,
                 const float* grad_slopes, int64_t gsl_stride, float* grad_pixels, void* stream) {
    if (!p || !pixels || !grad_slopes || !grad_pixels) return set_error(SHWFS_ERR_ARG, "null argument");
    if (n_frames <= 0 || n_frames > p->max_frames) return set_error(SHWFS_ERR_ARG, "n_frames outside [1, max_frames]");
    if (thr_mode != 0 && thr_mode != 1) return set_error(SHWFS_ERR_ARG, "thr_mode must be 0 or 1");
    const int64_t frame = (int64_t)p->frame_h * p->frame_w;
    if (n_frames > 1 && (pix_stride < frame || gsl_stride < 2 * (int64_t)p->n_sub))
        return set_error(SHWFS_ERR_ARG, "frame stride too small");
    DeviceGuard dg(p->device);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int npp = p->np_sub * p->np_sub;
    const double* xw = p->d_w;
    const double* yw = p->d_w + npp;
    const double* mw = p->d_w + 2 * npp;
    constexpr int kSubPerBlock = 4;
    dim3 grid((p->n_sub + kSubPerBlock - 1) / kSubPerBlock, n_frames);
    hipLaunchKernelGGL(cog_moments_kernel, grid, dim3(kSubPerBlock * kWave), 0, st, pixels, pix_stride,
                       p->d_origin, p->d_idx, p->n_sub, p->np_sub, p->frame_w, xw, yw, mw, thr, thr_mode, p->d_mom);
    SHWFS_LAUNCH_CHECK();
    hipLaunchKernelGGL(cog_finalize_kernel, dim3(n_frames), dim3(1024), 0, st, p->d_mom, p->n_sub,
                       (float*)nullptr, (int64_t)0, (double*)nullptr, (double*)nullptr, p->d_factor);
    SHWFS_LAUNCH_CHECK();
    hipLaunchKernelGGL(fill_zero_kernel, dim3(256), dim3(256), 0, st, grad_pixels,
                       (int64_t)(n_frames - 1) * frame + frame);
    SHWFS_LAUNCH_CHECK();
    hipLaunchKernelGGL(cog_vjp_kernel, grid, dim3(kSubPerBlock * kWave), 0, st, pixels, pix_stride, p->d_origin,
                       p->d_idx, p->n_sub, p->np_sub, p->frame_w, xw, yw, mw, thr, thr_mode, p->d_mom,
                       p->d_factor, grad_slopes, gsl_stride, grad_pixels, frame);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

int shwfs_im_from_jacobian(int n_lenses, int np_sub, int ef_size, const float* jx, const float* jy, int n_modes,
                           const float* modes, int64_t mode_stride, const int32_t* slot, int n_valid, float scale,
                           float* im, void* stream) {
    if (!jx || !jy || !modes || !slot || !im) return set_error(SHWFS_ERR_ARG, "null argument");
    if (n_lenses <= 0 || np_sub <= 0 || n_modes <= 0 || n_valid <= 0 || ef_size != n_lenses * np_sub)
        return set_error(SHWFS_ERR_ARG, "inconsistent sizes");
    if (ef_size > 1024) return set_error(SHWFS_ERR_UNSUPPORTED, "pupil wider than 1024 px");
    if (mode_stride < (int64_t)ef_size * ef_size) return set_error(SHWFS_ERR_ARG, "mode stride too small");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int threads = (ef_size + 63) / 64 * 64;
    const int mpb = 8;
    dim3 grid(n_lenses, (n_modes + mpb - 1) / mpb);
#define IMJ(NPV) hipLaunchKernelGGL(im_from_jacobian_kernel<NPV>, grid, dim3(threads), 0, st, jx, jy, ef_size, \
                                    n_lenses, modes, mode_stride, n_modes, mpb, slot, n_valid, scale, im)
    switch (np_sub) {
        case 4: IMJ(4); break;
        case 8: IMJ(8); break;
        case 16: IMJ(16); break;
        default: return set_error(SHWFS_ERR_UNSUPPORTED, "interaction matrix kernel instantiated for np_sub 4, 8, 16");
    }
#undef IMJ
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

}  // extern "C"

struct shwfs_batch {
    int device = 0;
    int n_sensors = 0;
    int n_items = 0;
    int grid = 0;
    NgsArgs* d_args = nullptr;
    BatchItem* d_items = nullptr;
    double* d_amp = nullptr;          // [n_sensors][2]
    double* d_amp_part = nullptr;
    unsigned* d_counter = nullptr;
    unsigned long long* d_ts = nullptr;   // [1 + n_sensors]
};

extern "C" {

int shwfs_batch_destroy(shwfs_batch* b) {
    if (!b) return SHWFS_OK;
    DeviceGuard dg(b->device);
    (void)hipFree(b->d_args);
    (void)hipFree(b->d_items);
    (void)hipFree(b->d_amp);
    (void)hipFree(b->d_amp_part);
    (void)hipFree(b->d_counter);
    (void)hipFree(b->d_ts);
    delete b;
    return SHWFS_OK;
}

int shwfs_batch_create(int n_sensors, shwfs_plan* const* plans, const int32_t* n_frames,
                       const float* const* A, const float* const* phase, float* const* ccd, const double* S0,
                       int grid_blocks, int order, int device, shwfs_batch** out) {
    if (n_sensors <= 0 || !plans || !n_frames || !A || !phase || !ccd || !S0 || !out)
        return set_error(SHWFS_ERR_ARG, "null argument");
    if (order < 0 || order > 2) return set_error(SHWFS_ERR_ARG, "order must be 0, 1 or 2");
    std::vector<NgsArgs> args(n_sensors);
    std::vector<BatchItem> items;
    std::vector<int> cost(n_sensors);
    for (int s = 0; s < n_sensors; ++s) {
        const shwfs_plan* p = plans[s];
        if (!p || !A[s] || !phase[s] || !ccd[s]) return set_error(SHWFS_ERR_ARG, "null sensor entry");
        if (p->device != device) return set_error(SHWFS_ERR_ARG, "plan on another device");
        const shwfs_geom& g = p->g;
        int kind;
        if (!g.lgs && p->mfma && g.np_sub == 8 && g.spec_size == 16) kind = 0;
        else if (g.lgs && g.np_sub == 16 && g.spec_size == 32) kind = 1;
        else return set_error(SHWFS_ERR_UNSUPPORTED, "batch supports NGS MFMA np 8 and LGS np 16 plans");
        if (!g.parseval) return set_error(SHWFS_ERR_UNSUPPORTED, "batch needs Parseval-normalised sensors");
        if (g.lgs && !p->kernels_set) return set_error(SHWFS_ERR_ARG, "LGS plan without kernels");
        if (n_frames[s] <= 0) return set_error(SHWFS_ERR_ARG, "non-positive frame count");
        const int ts = (kind == 0) ? 4 * NgsMfmaShape<8>::SW : 8;
        const int tiles = (p->n_sub + ts - 1) / ts;
        cost[s] = kind == 0 ? 1 : 4;
        NgsArgs& a = args[s];
        a = NgsArgs{};
        a.A = A[s]; a.a_stride = 0; a.phase = phase[s]; a.ph_stride = (int64_t)g.ef_size * g.ef_size;
        a.ccd = ccd[s];
        a.ccd_stride = (int64_t)g.n_lenses * g.subap_npx * g.n_lenses * g.subap_npx;
        a.roww = p->d_roww; a.colw = p->d_colw; a.amp_fstep = 0; a.bpart = nullptr;
        a.kspec = reinterpret_cast<const float2*>(p->d_kspec);
        a.n_lenses = g.n_lenses; a.ef_size = g.ef_size; a.npx = g.subap_npx; a.npx_pad = p->npx_pad;
        a.n_sub = p->n_sub; a.n_tiles = tiles; a.parseval = 1;
        a.k = g.k_rad_per_nm; a.tilt = g.tilt_rad; a.a0 = g.tilt_a0; a.S0 = S0[s];
        a.pixel_area = g.pixel_area; a.norm_const = g.norm_const;
        for (int f = 0; f < n_frames[s]; ++f)
            for (int t = 0; t < tiles; ++t) items.push_back(BatchItem{s, f, t, kind});
    }
    if (order == 0)
        std::stable_sort(items.begin(), items.end(),
                         [&](const BatchItem& x, const BatchItem& y) { return cost[x.sensor] > cost[y.sensor]; });
    else if (order == 2) {   // round robin across sensors
        std::vector<std::vector<BatchItem>> per(n_sensors);
        for (auto& it : items) per[it.sensor].push_back(it);
        items.clear();
        for (size_t r = 0;; ++r) {
            bool any = false;
            for (int s = 0; s < n_sensors; ++s)
                if (r < per[s].size()) { items.push_back(per[s][r]); any = true; }
            if (!any) break;
        }
    }
    DeviceGuard dg(device);
    auto* b = new shwfs_batch();
    b->device = device;
    b->n_sensors = n_sensors;
    b->n_items = (int)items.size();
    b->grid = grid_blocks > 0 ? std::min(grid_blocks, b->n_items) : b->n_items;
    auto cleanup = [&](int code) { shwfs_batch_destroy(b); return code; };
    hipError_t e;
    const int amp_blocks = 64;
    if ((e = hipMalloc((void**)&b->d_args, n_sensors * sizeof(NgsArgs))) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_items, items.size() * sizeof(BatchItem))) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_amp, n_sensors * 2 * sizeof(double))) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_amp_part, n_sensors * amp_blocks * 2 * sizeof(double))) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_counter, n_sensors * sizeof(unsigned))) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_ts, (n_sensors + 1) * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipMemset(b->d_counter, 0, n_sensors * sizeof(unsigned))) != hipSuccess)
        return cleanup(set_error(SHWFS_ERR_HIP, std::string("batch alloc: ") + hipGetErrorString(e)));
    // the pupils are static for the batch: amplitude statistics once, here
    for (int s = 0; s < n_sensors; ++s) {
        const int64_t px = (int64_t)plans[s]->g.ef_size * plans[s]->g.ef_size;
        hipLaunchKernelGGL(amp_stats_kernel, dim3(amp_blocks, 1), dim3(256), 0, 0, A[s], (int64_t)0, px,
                           b->d_amp_part + (size_t)s * amp_blocks * 2, b->d_counter + s, b->d_amp + 2 * s);
        args[s].amp = b->d_amp + 2 * s;
    }
    if ((e = hipGetLastError()) != hipSuccess || (e = hipDeviceSynchronize()) != hipSuccess ||
        (e = hipMemcpy(b->d_args, args.data(), n_sensors * sizeof(NgsArgs), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(set_error(SHWFS_ERR_HIP, std::string("batch setup: ") + hipGetErrorString(e)));
    *out = b;
    return SHWFS_OK;
}

int shwfs_batch_forward(shwfs_batch* b, void* stream) {
    if (!b) return set_error(SHWFS_ERR_ARG, "null batch");
    DeviceGuard dg(b->device);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(batch_stamp_kernel, dim3(1), dim3(64), 0, st, b->d_ts, b->n_sensors);
    SHWFS_LAUNCH_CHECK();
    hipLaunchKernelGGL(sh_batch_kernel, dim3(b->grid), dim3(256), kBatchLdsBytes, st, b->d_args, b->d_items,
                       b->n_items, b->d_ts);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

int shwfs_batch_latencies(shwfs_batch* b, double* us, void* stream) {
    if (!b || !us) return set_error(SHWFS_ERR_ARG, "null argument");
    DeviceGuard dg(b->device);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    std::vector<unsigned long long> ts(b->n_sensors + 1);
    SHWFS_HIP_CHECK(hipMemcpyAsync(ts.data(), b->d_ts, ts.size() * sizeof(unsigned long long),
                                   hipMemcpyDeviceToHost, st));
    SHWFS_HIP_CHECK(hipStreamSynchronize(st));
    int rate_khz = 0;
    SHWFS_HIP_CHECK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, b->device));
    if (rate_khz <= 0) return set_error(SHWFS_ERR_HIP, "wall clock rate unavailable");
    for (int s = 0; s < b->n_sensors; ++s)
        us[s] = (ts[1 + s] >= ts[0]) ? (double)(ts[1 + s] - ts[0]) * 1e3 / rate_khz : -1.0;
    return SHWFS_OK;
}

int shwfs_batch_info(shwfs_batch* b, int32_t* n_items, int32_t* grid) {
    if (!b || !n_items || !grid) return set_error(SHWFS_ERR_ARG, "null argument");
    *n_items = b->n_items;
    *grid = b->grid;
    return SHWFS_OK;
}

}  // extern "C"
