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

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/shwfs.h"
#include "fft_reg.h"

// Scratch of the no-allocation interaction-matrix entry point (see scratch_of_current_device).
constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

using namespace shwfs;

// ============================================================================ error plumbing
namespace {
thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
}  // namespace

// error reporting shared with the library's other translation units
int shwfs_internal_set_error(int code, const char* msg) { return set_error(code, msg); }

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

// Block-wide FP64 sums of two values at once (one pair of barriers); `red` needs blockDim/32
// doubles.  Results valid in every thread, fixed summation order.
__device__ double2 block_sum2(double a, double b, double* red) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const int nw = (blockDim.x + kWave - 1) / kWave;
    __syncthreads();
    if (lane == 0) { red[2 * wid] = a; red[2 * wid + 1] = b; }
    __syncthreads();
    double sa = 0.0, sb = 0.0;
    for (int w = 0; w < nw; ++w) { sa += red[2 * w]; sb += red[2 * w + 1]; }
    return make_double2(sa, sb);
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
// With `scale_out`, the Parseval flux scale S0 pixel_area sum(A) / (norm_const sum(A^2)) of the
// frame is also written (as float, the factor the SH kernels apply), `scale_num` = S0 pixel_area.
// With `B` (resampled field) the square sum is taken over B instead (sum(A) stays the masked area
// of the input field, sh.py:426).
__global__ void amp_stats_kernel(const float* __restrict__ A, int64_t a_stride, int64_t n_px,
                                 double* __restrict__ part, unsigned* __restrict__ counter,
                                 double* __restrict__ out, double scale_num, double norm_const,
                                 float* __restrict__ scale_out, const float* __restrict__ B,
                                 int64_t b_stride, int64_t n_b) {
    __shared__ double red[32];
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
    if (B) {
        s2 = 0.0;
        const float* b = B + f * b_stride;
        for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n_b; q += (int64_t)gridDim.x * blockDim.x) {
            const double v = b[q];
            s2 += v * v;
        }
    }
    const double2 s12 = block_sum2(s1, s2, red);
    double* p = part + ((int64_t)f * gridDim.x + blockIdx.x) * 2;
    if (threadIdx.x == 0) {
        p[0] = s12.x;
        p[1] = s12.y;
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
        const double2 t12 = block_sum2(t1, t2, red);
        t1 = t12.x;
        t2 = t12.y;
        if (threadIdx.x == 0) {
            out[2 * f] = t1;
            out[2 * f + 1] = t2;
            if (scale_out) scale_out[f] = (float)(scale_num * t1 / (norm_const * t2));
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
    const float* roww;   // P: npx_pad x M (NGS), or [L lags][16 rows] complex, transposed (LGS)
    const float* colw;   // npx_pad x M   (Q)
    const double* amp;   // [frame][2]: sum(A), sum(A^2)     (frame index f * amp_fstep)
    int amp_fstep;
    const float* fscale; // Parseval flux scale per amplitude frame (frame index f * scale_fstep)
    int scale_fstep;
    double* bpart;       // [frame][n_parts]: partial sums of unscaled B   (non-Parseval mode)
    const float2* kspec; // LGS: [n_sub][L][L] kernel spectra on the lag grid (L = 2 np - 1)
    const int32_t* slot; // fused CoG: [n_sub] slope index of each subaperture, -1 if not valid
    const int32_t* order; // NGS MFMA, LGS: [n_sub] subaperture at each wave / tile position (nullptr: identity)
    double* mom;         // fused CoG: [frame][n_valid][4] flux, thresholded total, x and y sums
    double* tpart;       // fused CoG: [frame][n_parts][2] per-wave sums of flux and thresholded total
    const float* cogw;   // fused CoG: x weights [npx], then y weights [npx]
    float thr;           // fused CoG threshold
    int n_valid;         // fused CoG: number of valid subapertures
    int vec;             // pupil rows may be read as float4
    int store_vec;       // CCD rows may be written as float4
    int cog;             // fused CoG mode: 0 off, 1 subtract-and-clamp, 2 pedestal
    int n_parts;         // partial sums per frame in bpart
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
        if (t == 0) g.bpart[(int64_t)f * g.n_parts + tile] = bsum;
    }
}

// ---------------------------------------------------------------- NGS forward, MFMA binning (M = 16)
// One wavefront owns SW = 64/NP consecutive subapertures and its own LDS region; the waves of a
// block never synchronise with each other.
//   A: lane (q, y)  -> E along pupil row y of subaperture q, pruned FFT_16 over x   -> R[q][kx][y]
//   B: lane (q, kx) -> autocorrelation over y of column kx (2np-1 lag reals)        -> C^T[q][kx][j]
//      (the power spectrum S = A C along y, A the cosine/sine synthesis, so P S = (P A) C: the
//      row operator uploaded for this kernel is P~ = P A)
//   C: per subaperture, B = P~ C Q^T as 2 x 4 v_mfma_f32_16x16x4f32:
//        T = C^T P~^T with the K index permuted as k = 4*(lane/16) + step, so that each lane reads
//                     one float4 of C^T (an LDS row chunk) and one float4 of P~;
//        D = Q T      where T, left in the accumulator layout (lane l, reg r = T[4(l/16)+r][l%16]),
//                     is directly the B operand of step r under the same K permutation.
//      D = B^T, so lane l ends with the float4 B[Y = l%16][X = 4(l/16) .. 4(l/16)+3]: one row chunk,
//      written back over the consumed S^T chunk it came from.
//   D: coalesced CCD stores of whole subaperture row segments, and (fused ShSlopec) the FP64
//      centre-of-gravity moments of each subaperture from the same registers.
// R (np complex per row) and S^T / B (16 reals per row) share the 64 bytes of an LDS row.
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Split of four floats into bf16 pieces [hi(4) | lo(4)]: hi = x rounded to bf16, lo = the
// residual x - hi (exact in FP32) rounded to bf16, so |x - hi - lo| <= 2^-17 |x| (unbiased).
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned uintx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 split_bf16(floatx4 v) {
    const bf16x2 h01 = {(__bf16)v[0], (__bf16)v[1]};     // one packed conversion per pair
    const bf16x2 h23 = {(__bf16)v[2], (__bf16)v[3]};
    const unsigned u01 = __builtin_bit_cast(unsigned, h01), u23 = __builtin_bit_cast(unsigned, h23);
    const bf16x2 l01 = {(__bf16)(v[0] - __uint_as_float(u01 << 16)), (__bf16)(v[1] - __uint_as_float(u01 & 0xffff0000u))};
    const bf16x2 l23 = {(__bf16)(v[2] - __uint_as_float(u23 << 16)), (__bf16)(v[3] - __uint_as_float(u23 & 0xffff0000u))};
    const uintx4 r = {u01, u23, __builtin_bit_cast(unsigned, l01), __builtin_bit_cast(unsigned, l23)};
    return __builtin_bit_cast(bf16x8, r);
}

// [lo | hi] from [hi | lo]
__device__ __forceinline__ bf16x8 swap_halves(bf16x8 v) {
    return __builtin_shufflevector(v, v, 4, 5, 6, 7, 0, 1, 2, 3);
}

template <int NP>
struct NgsMfmaShape {
    static constexpr int M = 16;
    static constexpr int SW = 64 / NP;               // subapertures per wave
    static constexpr int QS = M * M + 16;            // LDS words per subaperture (16 rows x 16 + 16 pad)
    static constexpr int WAVE_F4 = SW * QS / 4;      // float4 per wave
    static constexpr int WPB = 4;                    // waves per block
    static constexpr int LDS_F4 = WPB * WAVE_F4;     // float4 per block
    static constexpr int RPL = (SW * M + 63) / 64;   // phase-B rows per lane
};

// Chunk swizzle of LDS row r: 4-word chunk c of row r lives at chunk c ^ lds_swz(r).  With the
// 16-word subaperture pad (272 = 16 mod 64) every access pattern of the kernel is free of bank
// conflicts for 32 and for 64 banks: 16-byte reads of one chunk of 8/16 consecutive rows, 8-byte
// column writes of 2/4 consecutive subapertures, 16-byte row-segment reads of 2/4 subapertures.
__host__ __device__ constexpr int lds_swz(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

// CCD store layout of the fast path (npx = 4, 8 or 16 with np = 8): lane -> (subaperture q, float4
// column x4) and CCD rows Y = h, h + YSTEP, ... of that subaperture.
template <int NP, int NPX>
struct NgsStoreShape {
    static constexpr int SW = 64 / NP;
    static constexpr int X4 = NPX > 0 ? NPX / 4 : 1;
    static constexpr int LPR = X4 * SW;                          // lanes per row group
    static constexpr int YSTEP = NPX > 0 ? 64 / LPR : 1;
    static constexpr int NY = NPX > 0 ? (NPX + YSTEP - 1) / YSTEP : 1;   // rows per lane
};

// max(v, 0) on the bit pattern (one integer max, no NaN canonicalisation): non-negative floats
// order like their int32 patterns and every negative float (sign bit set) maps to +0.
__device__ __forceinline__ float clamp0(float v) { return __int_as_float(max(__float_as_int(v), 0)); }

// Exchanges within each row of 16 lanes (DPP, no LDS): the value of lane i's partner i ^ 1, i ^ 2,
// 7 - i (within its half row) or 15 - i.  Summing v and its partner's v along these four steps
// leaves the row's total in every lane of the row.
constexpr int kDppXor1 = 0xB1;          // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;          // quad_perm [2, 3, 0, 1]
constexpr int kDppHalfMirror = 0x141;   // row_half_mirror
constexpr int kDppMirror = 0x140;       // row_mirror
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

constexpr int kYwWord = 256;   // first padding word of subaperture 0 in a wave's LDS region

// Hand-over between lanes of one wave through LDS: a wave's LDS operations execute in order, so
// wavefront-scope fences around a wave barrier (no s_barrier) are enough.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Subapertures owned by one wave: positions sub0 .. sub0 + nq - 1 of the subaperture order
// (row-major lenslet index, or the plan's lit-first order for a registered static pupil).
struct NgsWave {
    int sub0;     // first position
    int nq;       // how many of the SW exist
    int i0, j0;   // lenslet row and column of sub0 (identity order)
    const int32_t* order;
};

// Lenslet index of the wave's subaperture q (q >= nq: the last one, for address arithmetic only).
__device__ __forceinline__ int wave_sub(const NgsWave& wv, int q) {
    const int pos = wv.sub0 + min(q, wv.nq - 1);
    return wv.order ? wv.order[pos] : pos;
}

__device__ __forceinline__ void wave_sub_rc(const NgsWave& wv, int n, int q, int& i, int& j) {
    if (wv.order) {
        const int s = wave_sub(wv, q);
        i = s / n;
        j = s - i * n;
        return;
    }
    j = wv.j0 + q;
    i = wv.i0;
    while (j >= n) { j -= n; ++i; }
}

template <int NP>
__device__ __forceinline__ NgsWave ngs_wave(const NgsArgs& g, int wave_id) {
    NgsWave wv;
    wv.sub0 = wave_id * NgsMfmaShape<NP>::SW;
    wv.nq = max(0, min(NgsMfmaShape<NP>::SW, g.n_sub - wv.sub0));
    wv.i0 = wv.sub0 / g.n_lenses;
    wv.j0 = wv.sub0 - wv.i0 * g.n_lenses;
    wv.order = g.order;
    return wv;
}

// Pupil samples of lane (q, y): row y of subaperture q (A and phase).  Kept separate from the
// transform so that a caller can issue the next frame's loads before transforming this one.
// `off`: pupil offset of this lane's row segment, -1 for lanes without a subaperture.
template <int NP>
__device__ __forceinline__ void ngs_row_load(const float* base, int off, bool vec, float (&v)[NP]) {
    if (off >= 0) {
        const float* r = base + (unsigned)off;
        if constexpr (NP % 4 == 0) {
            if (vec) {
#pragma unroll
                for (int c = 0; c < NP / 4; ++c) {
                    const float4 q = reinterpret_cast<const float4*>(r)[c];
                    v[4 * c] = q.x; v[4 * c + 1] = q.y; v[4 * c + 2] = q.z; v[4 * c + 3] = q.w;
                }
                return;
            }
        }
#pragma unroll
        for (int x = 0; x < NP; ++x) v[x] = r[x];
    } else {
#pragma unroll
        for (int x = 0; x < NP; ++x) v[x] = 0.f;
    }
}

// Phase row of frame f (only where the lane's pupil row has a nonzero sample: elsewhere the field
// is zero whatever the phase) and the frame's flux scale (Parseval mode), loaded together so that
// their latency is hidden together.  Lanes that do not load keep their registers: the caller zeroes
// them once, and a lane's `lit` does not change across the frames of one call.
template <int NP, bool PARSEVAL>
__device__ __forceinline__ void ngs_phase_load(const NgsArgs& g, int off, int f, bool lit, float (&pv)[NP],
                                               float& fs) {
    fs = PARSEVAL ? g.fscale[f * g.scale_fstep] : 1.f;
    if (lit && off >= 0) {
        const float* r = g.phase + f * g.ph_stride + (unsigned)off;
        if constexpr (NP % 4 == 0) {
            if (g.vec) {
#pragma unroll
                for (int c = 0; c < NP / 4; ++c) {
                    const float4 q = reinterpret_cast<const float4*>(r)[c];
                    pv[4 * c] = q.x; pv[4 * c + 1] = q.y; pv[4 * c + 2] = q.z; pv[4 * c + 3] = q.w;
                }
                return;
            }
        }
#pragma unroll
        for (int x = 0; x < NP; ++x) pv[x] = r[x];
    }
}

__device__ __forceinline__ bool any_nonzero_row(const float* v, int n) {
    bool lit = false;
    for (int x = 0; x < n; ++x) lit |= v[x] != 0.f;
    return lit;
}

// Per-wave constants that do not change from frame to frame.
template <int NP, int NPX, int COG>
struct NgsWaveConst {
    floatx4 pv, qv;       // binning operator chunks of this lane (P row l%16, Q row l%16, chunk l/16)
    bf16x8 pb_hl, pb_lh;  // BF3: [P_hi | P_lo] and [P_lo | P_hi] of pv
    // fused CoG y weights live in the 16 padding words of subaperture 0 (never touched by the
    // transform): LDS words kYwWord .. kYwWord + npx - 1
    int cword;            // LDS word of this lane's phase-C chunk (without the subaperture offset)
    int ld_off;           // pupil offset of this lane's row segment (-1: no subaperture)
    int st_off;           // fast store layout: CCD offset of this lane's first float4
    float xw[4];          // fused CoG: x weights of this lane's 4 columns
    int slot;             // fused CoG: slope index of this lane's subaperture (-1: not valid)
};

// Transform one frame of the wave's subapertures and store it (plus CoG moments when COG != 0:
// 1 = subtract-and-clamp threshold, 2 = pedestal threshold, 3 = no threshold (clamp at zero),
// sh_slopec.py:380-387).
// (phv, fs_io): this frame's phase row and flux scale, replaced by (phn, fn) -- the next frame's,
// prefetched by the caller -- once the transform no longer needs them.
template <int NP, bool PARSEVAL, int NPX, int COG, bool BF3>
__device__ __forceinline__ void ngs_mfma_frame(const NgsArgs& g, const NgsWave& wv, int wave_id, int f,
                                               const float (&av)[NP], float (&phv)[NP], float& fs_io,
                                               const float (&phn)[NP], const float& fn, float* sw,
                                               const NgsWaveConst<NP, NPX, COG>& wc, float sq_sub,
                                               unsigned long long lit) {
    const float fs = fs_io;
    auto rotate = [&]() {
#pragma unroll
        for (int x = 0; x < NP; ++x) phv[x] = phn[x];
        fs_io = fn;
    };
    using Sh = NgsMfmaShape<NP>;
    constexpr int M = Sh::M;
    constexpr int SW = Sh::SW;
    constexpr int QS = Sh::QS;
    static_assert(NP <= 8, "M = 16 path needs np_sub <= 8");
    const int lane = threadIdx.x & 63;
    const int n = g.n_lenses;

    // A subaperture whose pupil samples are all zero has an all-zero image.  A wave without a lit
    // subaperture skips the transforms and stores zeros; the binning skips dark subapertures, whose
    // LDS chunk then keeps the (zero) autocorrelation of phase B.  `lit`: bit q * NP + y set when
    // row y of subaperture q has a nonzero pupil sample (ngs_lit_rows).
    constexpr unsigned long long kSubBits = (NP == 64) ? ~0ull : ((1ull << NP) - 1);
    const bool dark = lit == 0;
    // what a dark wave stores: 0 times the frame's scale, so that a frame without any light gives
    // the reference's 0/0 (NaN) normalisation exactly as the full transform would
    const float dark_value = PARSEVAL ? 0.f * fs : 0.f;

    if (!dark) {
    {   // A: row transforms of the field
        const int q = lane / NP, y = lane - (lane / NP) * NP;
        const double k_rev = g.k * (1.0 / kTwoPi), t_rev = g.tilt * (1.0 / kTwoPi);
        const double cy = t_rev * (2.0 * g.a0 + y);
        float2 e[NP];
#pragma unroll
        for (int x = 0; x < NP; ++x) {
            double rv = fma((double)phv[x], k_rev, -fma(t_rev, (double)x, cy));
            rv -= rint(rv);
            const float rf = (float)rv;
            e[x] = make_float2(av[x] * __builtin_amdgcn_cosf(rf), av[x] * __builtin_amdgcn_sinf(rf));
        }
        float2 R[M];
        fft_pruned<M, NP, false>(e, R);
        if (q < SW) {
            const int base = q * QS + 2 * (y & 1);
            int cw[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) cw[s] = base + (((y >> 1) ^ s) << 2);
#pragma unroll
            for (int kx = 0; kx < M; ++kx)
                *reinterpret_cast<float2*>(sw + cw[lds_swz(kx)] + (kx << 4)) = R[kx];
        }
    }
    wave_lds_sync();

#pragma unroll 1
    for (int it = 0; it < Sh::RPL; ++it) {   // B: column transforms and power spectrum
        const int rho = lane + 64 * it;
        if ((SW * M) % 64 != 0 && rho >= SW * M) break;
        const int q = rho >> 4, kx = rho & 15;
        const int rowb = (q * QS + (kx << 4)) | (lds_swz(kx) << 2);
        constexpr int NC = (NP + 1) / 2;          // float4 chunks holding the np complex samples
        float2 col[2 * NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const float4 v = *reinterpret_cast<const float4*>(sw + (rowb ^ (c << 2)));
            col[2 * c] = make_float2(v.x, v.y);
            col[2 * c + 1] = make_float2(v.z, v.w);
        }
        // autocorrelation along y of column kx: C[d] = sum_y R[y] conj(R[y-d]), d = 0 .. np-1.
        // |FFT_16 over y|^2 = C[0] + 2 Re sum_d C[d] W^(ky d) is linear in these 2np-1 reals, so the
        // cosine/sine synthesis is folded into the row operator (P~ = P A on the host).
        float c[M];
#pragma unroll
        for (int j = 0; j < M; ++j) c[j] = 0.f;
#pragma unroll
        for (int y = 0; y < NP; ++y) c[0] = fmaf(col[y].x, col[y].x, fmaf(col[y].y, col[y].y, c[0]));
#pragma unroll
        for (int d = 1; d < NP; ++d) {
            float cr = 0.f, ci = 0.f;
#pragma unroll
            for (int y = d; y < NP; ++y) {
                cr = fmaf(col[y].x, col[y - d].x, fmaf(col[y].y, col[y - d].y, cr));
                ci = fmaf(col[y].y, col[y - d].x, fmaf(-col[y].x, col[y - d].y, ci));
            }
            c[d] = cr;
            c[NP - 1 + d] = ci;
        }
#pragma unroll
        for (int q4 = 0; q4 < M / 4; ++q4)
            *reinterpret_cast<float4*>(sw + (rowb ^ (q4 << 2))) =
                make_float4(c[4 * q4], c[4 * q4 + 1], c[4 * q4 + 2], c[4 * q4 + 3]);
    }
    wave_lds_sync();

    // C: binning on the matrix cores; the frame's flux scale is folded into Q (Parseval mode)
    floatx4 qv = wc.qv;
    if constexpr (PARSEVAL) qv *= fs;
    bf16x8 qa_hl, qa_lh;                 // BF3: [Q_hi | Q_lo], [Q_lo | Q_hi] of the scaled Q chunk
    if constexpr (BF3) {
        qa_hl = split_bf16(qv);
        qa_lh = swap_halves(qa_hl);
    }
    auto bin_one = [&](int q) {
        if (((lit >> (q * NP)) & kSubBits) == 0) return;       // dark subaperture: chunk stays zero
        floatx4* chunk = reinterpret_cast<floatx4*>(sw + q * QS + wc.cword);
        const floatx4 sv = *chunk;
        if constexpr (BF3) {
            // FP32-level products from bf16 pieces x = x_hi + x_lo: the K slots of a lane hold
            // [hi(4) | lo(4)] of its four j (or kx) values, so [x_hi | x_lo] . [y_hi | y_lo] and
            // [x_hi | x_lo] . [y_lo | y_hi] (two 16x16x32 products) give all four cross terms of
            // (x_hi + x_lo)(y_hi + y_lo); what remains is the 2^-17 representation error.
            const bf16x8 a = split_bf16(sv);                         // [C_hi | C_lo]
            floatx4 t = {0.f, 0.f, 0.f, 0.f};
            t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, wc.pb_hl, t, 0, 0, 0);   // C_hi P_hi + C_lo P_lo
            t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, wc.pb_lh, t, 0, 0, 0);   // C_hi P_lo + C_lo P_hi
            const bf16x8 b = split_bf16(t);                          // [T_hi | T_lo]
            floatx4 d = {0.f, 0.f, 0.f, 0.f};
            d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(qa_hl, b, d, 0, 0, 0);      // Q_hi T_hi + Q_lo T_lo
            d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(qa_lh, b, d, 0, 0, 0);      // Q_lo T_hi + Q_hi T_lo
            *chunk = d;
            return;
        }
        floatx4 t = {0.f, 0.f, 0.f, 0.f};
        t = __builtin_amdgcn_mfma_f32_16x16x4f32(sv[0], wc.pv[0], t, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x4f32(sv[1], wc.pv[1], t, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x4f32(sv[2], wc.pv[2], t, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x4f32(sv[3], wc.pv[3], t, 0, 0, 0);
        floatx4 d = {0.f, 0.f, 0.f, 0.f};
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(qv[0], t[0], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(qv[1], t[1], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(qv[2], t[2], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(qv[3], t[3], d, 0, 0, 0);
        *chunk = d;
    };
    if (wv.nq == SW) {
#pragma unroll
        for (int q = 0; q < SW; ++q) bin_one(q);
    } else {
        for (int q = 0; q < wv.nq; ++q) bin_one(q);
    }
    rotate();
    wave_lds_sync();
    } else {
        rotate();
    }   // !dark

    // D: stores (+ moments)
    float* ccdf = g.ccd + f * g.ccd_stride;
    const int W = n * g.npx;
    double bsum = 0.0;
    if constexpr (NPX > 0) {
        using St = NgsStoreShape<NP, NPX>;
        constexpr int X4 = St::X4, LPR = St::LPR, YSTEP = St::YSTEP, NY = St::NY;
        static_assert(NPX % 4 == 0 && 64 % LPR == 0, "row groups must tile the wave");
        const int x4 = lane % X4, q = (lane / X4) % SW, h = lane / LPR;
        const bool live = q < wv.nq;
        (void)x4; (void)W;
        float* ccdw = ccdf + wc.st_off;
        // this lane's 4 x NY pixels: thresholded column sums and y moment (FP32 partials)
        float l_y = 0.f, cw0 = 0.f, cw1 = 0.f, cw2 = 0.f, cw3 = 0.f;
        if (live) {
            if (dark) {                  // wave-uniform: no transform ran, every pixel is 0 x scale
                const floatx4 o = {dark_value, dark_value, dark_value, dark_value};
#pragma unroll
                for (int t = 0; t < NY; ++t) {
                    if (NPX % YSTEP != 0 && h + YSTEP * t >= NPX) break;
                    __builtin_nontemporal_store(o, reinterpret_cast<floatx4*>(ccdw + (unsigned)(YSTEP * t * W)));
                }
                if constexpr (!PARSEVAL) bsum = 0.0;
                if constexpr (COG != 0) cw0 = cw1 = cw2 = cw3 = l_y = dark_value;
            } else {
#pragma unroll
                for (int t = 0; t < NY; ++t) {
                    const int Y = h + YSTEP * t;
                    if (NPX % YSTEP != 0 && Y >= NPX) break;
                    const float4 v = *reinterpret_cast<const float4*>(sw + q * QS + (Y << 4) + ((x4 ^ lds_swz(Y)) << 2));
                    const floatx4 o = {v.x, v.y, v.z, v.w};
                    __builtin_nontemporal_store(o, reinterpret_cast<floatx4*>(ccdw + (unsigned)(YSTEP * t * W)));
                    if constexpr (!PARSEVAL) bsum += (double)((v.x + v.y) + (v.z + v.w));
                    if constexpr (COG != 0) {
                        float w0, w1, w2, w3;
                        if constexpr (COG == 1) {          // subtract the threshold, clamp at zero
                            w0 = clamp0(v.x - g.thr); w1 = clamp0(v.y - g.thr);
                            w2 = clamp0(v.z - g.thr); w3 = clamp0(v.w - g.thr);
                        } else if constexpr (COG == 2) {   // pedestal
                            w0 = v.x < g.thr ? 0.f : v.x; w1 = v.y < g.thr ? 0.f : v.y;
                            w2 = v.z < g.thr ? 0.f : v.z; w3 = v.w < g.thr ? 0.f : v.w;
                        } else {
                            // no threshold: the reference clamps at zero (sh_slopec.py:382-386), a no-op on
                            // the exact image, which is a sum of squared moduli; the FP32 image's negative
                            // pixels are rounding residues far below its own error, so they are kept
                            w0 = v.x; w1 = v.y; w2 = v.z; w3 = v.w;
                        }
                        const float yv = sw[kYwWord + Y];   // y weight of row Y (kept in LDS padding)
                        l_y = fmaf((w0 + w1) + (w2 + w3), yv, l_y);
                        cw0 += w0; cw1 += w1; cw2 += w2; cw3 += w3;
                    }
                }
            }
        }
        if constexpr (COG != 0) {
            // flux (pre-threshold subaperture sum, sh_slopec.py:365) from the Parseval identity of this
            // geometry: the subaperture's image sums to norm_const x its sum(A^2), times the frame scale
            const double m_f = (double)sq_sub * ((double)fs * g.norm_const);
            // FP32 partials of 4 x NY pixels per lane; FP64 across lanes
            const float l_x = fmaf(cw3, wc.xw[3], fmaf(cw2, wc.xw[2], fmaf(cw1, wc.xw[1], cw0 * wc.xw[0])));
            // the total comes from the same pixels as the x and y sums (sh_slopec.py:392), so that
            // errors common to the subaperture's pixels cancel in the slopes
            double m_x = l_x, m_y = l_y, m_t = (double)((cw0 + cw1) + (cw2 + cw3));
            // the column groups of one row: lanes x4 ^ 1 and x4 ^ 2 of the same quad (DPP)
            if constexpr (X4 > 1) {
                m_t += dpp_d<kDppXor1>(m_t); m_x += dpp_d<kDppXor1>(m_x); m_y += dpp_d<kDppXor1>(m_y);
            }
            if constexpr (X4 > 2) {
                static_assert(X4 == 4, "at most 4 column groups");
                m_t += dpp_d<kDppXor2>(m_t); m_x += dpp_d<kDppXor2>(m_x); m_y += dpp_d<kDppXor2>(m_y);
            }
#pragma unroll
            for (int o = LPR; o < 64; o <<= 1) {
                m_t += __shfl_xor(m_t, o, kWave);
                m_x += __shfl_xor(m_x, o, kWave); m_y += __shfl_xor(m_y, o, kWave);
            }
            const bool writer = live && lane < LPR && x4 == 0 && wc.slot >= 0;
            if (writer) {
                double* m = g.mom + ((int64_t)f * g.n_valid + wc.slot) * 4;
                *reinterpret_cast<double2*>(m) = make_double2(m_f, m_t);
                *reinterpret_cast<double2*>(m + 2) = make_double2(m_x, m_y);
            }
            // the wave's share of the frame totals (flux, thresholded total), fixed order
            double p_f = writer ? m_f : 0.0, p_t = writer ? m_t : 0.0;
#pragma unroll
            for (int o = X4; o < LPR; o <<= 1) {
                p_f += __shfl_xor(p_f, o, kWave);
                p_t += __shfl_xor(p_t, o, kWave);
            }
            if (lane == 0)
                *reinterpret_cast<double2*>(g.tpart + ((int64_t)f * g.n_parts + wave_id) * 2) = make_double2(p_f, p_t);
        }
    } else {
        const int npx = g.npx;
        const int per = wv.nq * npx * npx;
        for (int e = lane; e < per; e += 64) {
            const int X = e % npx, t2 = e / npx;
            const int q = t2 % wv.nq, Y = t2 / wv.nq;
            int i, j;
            wave_sub_rc(wv, n, q, i, j);
            const float v = dark ? dark_value : sw[q * QS + (Y << 4) + (((X >> 2) ^ lds_swz(Y)) << 2) + (X & 3)];
            if constexpr (!PARSEVAL) bsum += (double)v;
            ccdf[(i * npx + Y) * W + j * npx + X] = v;
        }
    }
    if constexpr (!PARSEVAL) {
        bsum = wave_sum(bsum);
        if (lane == 0) g.bpart[(int64_t)f * g.n_parts + wave_id] = bsum;
    }
}

template <int NP, int NPX, int COG>
__device__ __forceinline__ NgsWaveConst<NP, NPX, COG> ngs_wave_const(const NgsArgs& g, const NgsWave& wv,
                                                                    float* sw) {
    constexpr int M = NgsMfmaShape<NP>::M;
    const int lane = threadIdx.x & 63;
    const int i16 = lane & 15, g4 = lane >> 4;
    NgsWaveConst<NP, NPX, COG> wc;
    wc.pv = *reinterpret_cast<const floatx4*>(g.roww + i16 * M + 4 * g4);
    wc.qv = *reinterpret_cast<const floatx4*>(g.colw + i16 * M + 4 * g4);
    wc.cword = (i16 << 4) + ((g4 ^ lds_swz(i16)) << 2);
    wc.pb_hl = split_bf16(wc.pv);
    wc.pb_lh = swap_halves(wc.pb_hl);
    wc.slot = -1;
    {
        const int q = lane / NP, y = lane - (lane / NP) * NP;
        wc.ld_off = -1;
        if (q < wv.nq) {
            int i, j;
            wave_sub_rc(wv, g.n_lenses, q, i, j);
            wc.ld_off = (i * NP + y) * g.ef_size + j * NP;
        }
    }
    wc.st_off = 0;
    if constexpr (NPX > 0) {
        using St = NgsStoreShape<NP, NPX>;
        const int x4 = lane % St::X4, q = (lane / St::X4) % St::SW, h = lane / St::LPR;
        int i, j;
        wave_sub_rc(wv, g.n_lenses, q, i, j);
        wc.st_off = (i * NPX + h) * (g.n_lenses * NPX) + j * NPX + 4 * x4;
    }
    if constexpr (COG != 0 && NPX > 0) {
        using St = NgsStoreShape<NP, NPX>;
        constexpr int SW = St::SW, X4 = St::X4, LPR = St::LPR, YSTEP = St::YSTEP, NY = St::NY;
        const int x4 = lane % X4, q = (lane / X4) % SW, h = lane / LPR;
#pragma unroll
        for (int c = 0; c < 4; ++c) wc.xw[c] = g.cogw[4 * x4 + c];
        if (lane < NPX) sw[kYwWord + lane] = g.cogw[NPX + lane];
        wave_lds_sync();
        (void)YSTEP; (void)NY; (void)h;
        if (q < wv.nq) wc.slot = g.slot[wave_sub(wv, q)];
    }
    return wc;
}

// Sum of A^2 over the pupil samples of a subaperture, delivered to the lanes of the CCD store layout
// (NgsStoreShape) that hold it: per pupil row in the load layout (lane (q, y)), across the NP rows of
// subaperture q, then fetched by the store lanes of q.  FP32 (exact for a 0/1 pupil).
template <int NP, int NPX>
__device__ __forceinline__ float subap_sq_store_layout(const float (&av)[NP]) {
    static_assert((NP & (NP - 1)) == 0, "rows of a subaperture are reduced with xor steps");
    float s = 0.f;
#pragma unroll
    for (int x = 0; x < NP; ++x) s = fmaf(av[x], av[x], s);
#pragma unroll
    for (int o = 1; o < NP; o <<= 1) s += __shfl_xor(s, o, kWave);
    using St = NgsStoreShape<NP, NPX>;
    const int q = ((threadIdx.x & 63) / St::X4) % St::SW;
    return __shfl(s, q * NP, kWave);
}

// s_waitcnt immediate that waits for every vector memory operation (vmcnt 0) and leaves the export
// and LDS / scalar counters at their maxima (no wait on them).
constexpr int kWaitVmcnt0 = (0x7 << 4) | (0xF << 8);

// Rows of the wave's subapertures with a nonzero pupil sample (bit q * NP + y: row y of q).
template <int NP>
__device__ __forceinline__ unsigned long long ngs_lit_rows(const float (&av)[NP]) {
    bool lit_lane = false;
#pragma unroll
    for (int x = 0; x < NP; ++x) lit_lane |= (av[x] != 0.f);
    return __ballot(lit_lane);
}

// Frames [f0, f1) of the wave's subapertures; the next frame's phase loads are issued before the
// current frame is transformed, so HBM latency overlaps the FFT / MFMA work.  Vector memory
// operations are counted in order (loads and stores alike), so a wait for a register loaded before
// other operations also waits for them: the loop is entered with every load complete, and the
// prefetched phase is moved into the current-frame registers after the binning, before the frame's
// stores are issued (a move after the stores would wait for them, one at the top of the next frame
// for the loads of the frame after it).
template <int NP, bool PARSEVAL, int NPX, int COG, bool BF3>
__device__ __forceinline__ void ngs_mfma_frames(const NgsArgs& g, int wave_id, int f0, int f1, float* sw) {
    const NgsWave wv = ngs_wave<NP>(g, wave_id);
    if (wv.nq == 0) {                     // past the last subaperture: empty partial sums only
        if ((threadIdx.x & 63) == 0)
            for (int f = f0; f < f1; ++f) {
                if constexpr (!PARSEVAL) g.bpart[(int64_t)f * g.n_parts + wave_id] = 0.0;
                if constexpr (COG != 0)
                    *reinterpret_cast<double2*>(g.tpart + ((int64_t)f * g.n_parts + wave_id) * 2) = make_double2(0.0, 0.0);
            }
        return;
    }
    const NgsWaveConst<NP, NPX, COG> wc = ngs_wave_const<NP, NPX, COG>(g, wv, sw);
    float av[NP], pv[NP], pn[NP], fs, fn = 1.f;
    ngs_row_load<NP>(g.A + f0 * g.a_stride, wc.ld_off, g.vec, av);
    // static pupil (a_stride 0): the amplitude is loaded once, and a lane whose pupil row has no
    // nonzero sample never loads its phase (the field there is zero whatever the phase; its phase
    // registers stay zero); a per-frame pupil loads every row, its amplitude at the top of each
    // frame (only the phase is prefetched a frame ahead)
    const bool static_a = g.a_stride == 0;
    const bool lit = !static_a || any_nonzero_row(av, NP);
#pragma unroll
    for (int x = 0; x < NP; ++x) { pv[x] = 0.f; pn[x] = 0.f; }
    ngs_phase_load<NP, PARSEVAL>(g, wc.ld_off, f0, lit, pv, fs);
    // fused CoG: sum(A^2) of each store lane's subaperture (once for a static pupil, else per frame)
    float sq = 0.f;
    if constexpr (COG != 0 && NPX > 0) sq = subap_sq_store_layout<NP, NPX>(av);
    unsigned long long litm = ngs_lit_rows<NP>(av);
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    for (int f = f0; f < f1; ++f) {
        if (!static_a && f > f0) {
            ngs_row_load<NP>(g.A + f * g.a_stride, wc.ld_off, g.vec, av);
            litm = ngs_lit_rows<NP>(av);
            if constexpr (COG != 0 && NPX > 0) sq = subap_sq_store_layout<NP, NPX>(av);
        }
        if (f + 1 < f1) ngs_phase_load<NP, PARSEVAL>(g, wc.ld_off, f + 1, lit, pn, fn);
        if (f > f0) wave_lds_sync();
        // transforms frame f from (pv, fs) and moves (pn, fn) into them before its stores
        ngs_mfma_frame<NP, PARSEVAL, NPX, COG, BF3>(g, wv, wave_id, f, av, pv, fs, pn, fn, sw, wc, sq, litm);
    }
}

// grid = (n_tiles * WPB, ceil(n_frames / fpb)); block = one wave (the waves share nothing, so a
// wave of dark subapertures that ends early frees its slot and its LDS for the next one).
template <int NP, bool PARSEVAL, int NPX, int COG, bool BF3>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(NP == 8 && !(BF3 && (COG == 1 || COG == 2)) ? 4 : 3, 8)))
sh_ngs_mfma_kernel(NgsArgs g, int n_frames, int fpb) {
    using Sh = NgsMfmaShape<NP>;
    __shared__ float4 lds_wave[Sh::WAVE_F4];
    const int f0 = blockIdx.y * fpb;
    const int f1 = min(n_frames, f0 + fpb);
    ngs_mfma_frames<NP, PARSEVAL, NPX, COG, BF3>(g, blockIdx.x, f0, f1, reinterpret_cast<float*>(lds_wave));
}

// ---------------------------------------------------------------- general path (any np, M)
#include "sh_general.h"

// ---------------------------------------------------------------- LGS forward (sh.py:381-394)
// psf = Re IFFT_N(FFT_N(|FFT_N(E)|^2) * K).  With D[m] = (N^2/M^2) FFT_M(S)[m mod M] K[m] on the
// lag grid |m| <= np-1 (S = |FFT_M(E)|^2, M >= 2np-1), the binned image is B = Re(P D Q^T) where
// P, Q (npx x L, complex) already carry N^2/M^2, the rebin, the cut and the mask.
// One block = TS subapertures, 64 x WAVES threads (WAVES = TS / (2 SPN): two 16-column blocks per wave
// in the matrix-core phases):
//   1:  task (s, x)       column FFT over y of the field                      -> buf[s][ky][x]
//   2:  task (s, ky)      lag row of |FFT_x(.)|^2 by direct autocorrelation     -> buf[s][lx][ky]
//   3a: task (s, lx)      forward FFT over ky, D = F * K_s, in place            -> buf[s][lx][my]
//   3b: MFMA              V[Y][lx] = sum_my P[Y][my] D[my][lx] (FP32 16x16x4)   -> buf[s][Y][lx]
//   4:  MFMA              B[Y][X] = Re sum_lx V[Y][lx] Q[X][lx] (FP32 16x16x4)  -> CCD
// Every LDS row is padded by one complex element, so lanes that walk different rows of one
// array hit different banks (row lengths NP, M and LH are multiples of 32 dwords otherwise).
template <int NP, int M, int TS>
struct LgsShape {
    static constexpr int L = 2 * NP - 1;
    static constexpr int LH = NP;                 // lag columns m_x = 0 .. NP-1 (Hermitian half)
    static constexpr int R1 = NP + 1, R2 = M + 1, R3 = LH + 1;   // padded row lengths
    static constexpr int E1 = M * R1, E2 = LH * R2, E3 = 16 * R3;
    static constexpr int LDS_F2 = TS * (E1 > E2 ? (E1 > E3 ? E1 : E3) : (E2 > E3 ? E2 : E3));
    static constexpr int SPN = 16 / LH;           // subapertures per 16-column block of the MFMA phases
    static constexpr int WAVES = TS / (2 * SPN);  // two column blocks per wave
    static constexpr int BT = 64 * WAVES;         // threads per block
    // fused CoG: LDS after the tile buffer for the column square sums of the pupil samples (one
    // per phase-1 task)
    static constexpr int COG_BYTES = TS * NP * 4;
};

// The lag-domain image D = C K is Hermitian on the (m_y, m_x) grid (C is an autocorrelation, K
// the spectrum of a real kernel), and so are the binning operators (transforms of real ones):
// Re sum_{m_y, m_x} P[Y][m_y] D[m_y][m_x] Q[X][m_x] only needs the columns m_x >= 0, the
// m_x > 0 ones counted twice.  Phases 2-4 therefore carry LH = NP lag columns instead of 2NP-1.
//
// COG (1 subtracted threshold, 2 pedestal, 3 no threshold): the ShSlopec centre of gravity
// (sh_slopec.py:365-404) of the tile's subapertures from the
// binned pixels of phase 4 while they are in registers: per subaperture the thresholded total
// and the x / y moments (FP32 over 32 pixels, FP64 across lanes) and the Parseval flux
// (scale x norm_const x sum(A^2) of the subaperture), into g.mom at the subaperture's slot, plus
// its share of the frame totals into g.tpart (one pair per tile position: g.n_parts = tiles x TS);
// cog_finalize_parts_kernel turns them into slopes.  `buf` then extends Sh::COG_BYTES past the
// tile buffer.
template <int NP, int M, int TS, int COG = 0>
__device__ __forceinline__ void lgs_tile(const NgsArgs& g, int tile, int f, float2* buf, double* red) {
    constexpr int L = 2 * NP - 1;
    using Sh = LgsShape<NP, M, TS>;
    constexpr int LH = Sh::LH, R1 = Sh::R1, R2 = Sh::R2, BT = Sh::BT, WAVES = Sh::WAVES;
    static_assert(L <= M, "lag grid must fit the spectrum grid");
    const int t = threadIdx.x;
    const int n = g.n_lenses;
    const int sub0 = tile * TS;
    const float2* Qc = reinterpret_cast<const float2*>(g.colw);
    float* cog_sq = reinterpret_cast<float*>(buf + Sh::LDS_F2);        // [TS * NP]

    // tile position s -> lenslet index (the plan's lit-first order for a registered static pupil)
    auto lenslet = [&](int s) { const int pos = sub0 + s; return g.order ? g.order[pos] : pos; };

    // phase 1a: pupil samples of the task's column (one task per thread); a tile without a lit
    // sample has a zero image and stores it without transforming
    static_assert(TS * NP <= BT, "one phase-1 task per thread");
    const int s1 = t / NP, x1 = t - (t / NP) * NP;
    const bool task1 = t < TS * NP && sub0 + s1 < g.n_sub;
    int64_t off1 = 0;
    float av[NP];
    bool lit = false;
#pragma unroll
    for (int y = 0; y < NP; ++y) av[y] = 0.f;
    if (task1) {
        const int sub = lenslet(s1);
        const int i = sub / n, j = sub - (sub / n) * n;
        off1 = (int64_t)(i * NP) * g.ef_size + (int64_t)j * NP + x1;
        const float* ar = g.A + f * g.a_stride + off1;
#pragma unroll
        for (int y = 0; y < NP; ++y) {
            av[y] = ar[(int64_t)y * g.ef_size];
            lit |= av[y] != 0.f;
        }
    }
    if constexpr (COG != 0) {
        if (t < TS * NP) {
            float q = 0.f;
#pragma unroll
            for (int y = 0; y < NP; ++y) q = fmaf(av[y], av[y], q);
            cog_sq[t] = q;                            // visible after the barrier below
        }
    }
    double scale = 1.0;                               // Parseval flux scale of the frame
    if (g.parseval) {
        const double* am = g.amp + (int64_t)f * g.amp_fstep;
        scale = g.S0 * g.pixel_area * am[0] / (g.norm_const * am[1]);
    }
    if (!__syncthreads_or(lit)) {
        if constexpr (COG != 0) {
            // every pixel of the tile is 0 x the scale: so are its moments and flux
            const double zd = 0.0 * scale;
            if (t < TS) {
                const int slot = sub0 + t < g.n_sub ? g.slot[lenslet(t)] : -1;
                if (slot >= 0) {
                    double* m = g.mom + ((int64_t)f * g.n_valid + slot) * 4;
                    *reinterpret_cast<double2*>(m) = make_double2(zd, zd);
                    *reinterpret_cast<double2*>(m + 2) = make_double2(zd, zd);
                }
                const double z = slot >= 0 ? zd : 0.0;
                *reinterpret_cast<double2*>(g.tpart + ((int64_t)f * g.n_parts + tile * TS + t) * 2) = make_double2(z, z);
            }
        }
        // 0 times the scale: a frame without any light gives the reference's 0/0 (NaN) as the
        // full transform would
        const float z = (float)(0.0 * scale);
        const int W = n * g.npx;
        if (g.store_vec && (g.npx & 3) == 0) {
            const int q4 = g.npx / 4, per = g.npx * q4;
            for (int e = t; e < TS * per; e += BT) {
                const int s = e / per, r = e - (e / per) * per;
                if (sub0 + s >= g.n_sub) break;
                const int sub = lenslet(s);
                const int i = sub / n, j = sub - (sub / n) * n, Y = r / q4, X4 = r - (r / q4) * q4;
                float* dst = g.ccd + f * g.ccd_stride + (int64_t)(i * g.npx + Y) * W + (int64_t)j * g.npx + 4 * X4;
                *reinterpret_cast<floatx4*>(dst) = floatx4{z, z, z, z};
            }
        } else {
            const int per = g.npx * g.npx;
            for (int e = t; e < TS * per; e += BT) {
                const int s = e / per, r = e - (e / per) * per;
                if (sub0 + s >= g.n_sub) break;
                const int sub = lenslet(s);
                const int i = sub / n, j = sub - (sub / n) * n, Y = r / g.npx, X = r - (r / g.npx) * g.npx;
                g.ccd[f * g.ccd_stride + (int64_t)(i * g.npx + Y) * W + (int64_t)j * g.npx + X] = z;
            }
        }
        if (!g.parseval && t == 0) g.bpart[(int64_t)f * g.n_parts + tile] = 0.0;   // its share of sum(B)
        return;
    }

    // phase 1b: field and column FFT over y
    if (t < TS * NP) {
        float2 e[NP];
        if (task1) {
            const float* pr = g.phase + f * g.ph_stride + off1;
            const double k_rev = g.k * (1.0 / kTwoPi), tilt_rev = g.tilt * (1.0 / kTwoPi);
            const double ax = g.a0 + x1;
#pragma unroll
            for (int y = 0; y < NP; ++y)
                e[y] = field_sample_rev(av[y], pr[(int64_t)y * g.ef_size], k_rev, tilt_rev, ax + g.a0 + y);
        } else {
#pragma unroll
            for (int y = 0; y < NP; ++y) e[y] = make_float2(0.f, 0.f);
        }
        float2 G[M];
        fft_pruned<M, NP, false>(e, G);
#pragma unroll
        for (int ky = 0; ky < M; ++ky) buf[(s1 * M + ky) * R1 + x1] = G[ky];
    }
    __syncthreads();

    // phase 2 (read rows into registers, then overwrite the buffer with lag rows m_x >= 0)
    constexpr int T2 = (TS * M + BT - 1) / BT;
    float2 R[T2][LH];
#pragma unroll
    for (int h = 0; h < T2; ++h) {
        const int task = t + BT * h;
        if (task < TS * M) {
            const int s = task / M, ky = task - (task / M) * M;
            float2 row[NP];
#pragma unroll
            for (int x = 0; x < NP; ++x) row[x] = buf[(s * M + ky) * R1 + x];
            // FFT_M(|FFT_M(row)|^2)[m] = M sum_x row[x] conj(row[x + m]) (linear for |m| < NP since
            // row is zero beyond NP): the lags m_x = 0 .. NP-1 directly, NP (NP + 1) / 2 products;
            // the factor M is carried by the row operator, and lag 0 is real
            {
                float re = 0.f;
#pragma unroll
                for (int x = 0; x < NP; ++x) re = fmaf(row[x].x, row[x].x, fmaf(row[x].y, row[x].y, re));
                R[h][0] = make_float2(re, 0.f);
            }
#pragma unroll
            for (int lh = 1; lh < LH; ++lh) {
                float re = 0.f, im = 0.f;
#pragma unroll
                for (int x = 0; x + lh < NP; ++x) {
                    re = fmaf(row[x].x, row[x + lh].x, fmaf(row[x].y, row[x + lh].y, re));
                    im = fmaf(row[x].y, row[x + lh].x, fmaf(-row[x].x, row[x + lh].y, im));
                }
                R[h][lh] = make_float2(re, im);                             // m_x = lh
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < T2; ++h) {
        const int task = t + BT * h;
        if (task < TS * M) {
            const int s = task / M, ky = task - (task / M) * M;
#pragma unroll
            for (int lh = 0; lh < LH; ++lh) buf[(s * LH + lh) * R2 + ky] = R[h][lh];
        }
    }
    __syncthreads();

    // phase 3a: column transforms over ky and the kernel multiply, written back in place
    for (int task = t; task < TS * LH; task += BT) {
        const int s = task / LH, lh = task - (task / LH) * LH;
        if (sub0 + s >= g.n_sub) continue;
        const int sub = lenslet(s);
        float2* col = buf + (s * LH + lh) * R2;
        float2 F[M];
#pragma unroll
        for (int ky = 0; ky < M; ++ky) F[ky] = col[ky];
        fft_radix2<M, false>(F);
        const float2* kr = g.kspec + (int64_t)sub * L * L + lh + NP - 1;   // lag column m_x = lh
#pragma unroll
        for (int mi = 0; mi < L; ++mi) col[mi] = cmul(F[(mi - (NP - 1) + M) % M], kr[mi * L]);
    }
    __syncthreads();

    // phase 3b: left binning operator V = P~ D on the matrix cores (FP32 16x16x4, K = lags m_y).
    // A = P~ (16 rows Y x 4 lags per step: lane l holds P~[l % 16][4 ks + l / 16], kept in
    // registers for the tile), B = D (4 lags x 16 columns: SPN = 16 / LH subapertures side by
    // side, lag column m_x = column % LH), V[Y][n] out in lane n + 16 (Y / 4), register Y % 4.
    // Three real products per K step (Gauss / Karatsuba): T1 = Re P~ Re D, T2 = Im P~ Im D,
    // T3 = (Re P~ + Im P~)(Re D + Im D); Re V = T1 - T2, Im V = T3 - T1 - T2.
    {
        constexpr int SPN = Sh::SPN;                 // subapertures per 16-column block
        constexpr int KS = (L + 3) / 4;              // K steps (lags beyond L: P~ = 0)
        static_assert(TS == 2 * WAVES * SPN, "two column blocks per wave");
        static_assert(4 * KS <= M, "K steps stay inside the written part of an LDS row");
        const int lane = t & 63, wave = t >> 6;
        const int n16 = lane & 15, k4 = lane >> 4;
        const float2* Pt = reinterpret_cast<const float2*>(g.roww);       // [m_y][16 rows] complex
        float pr[KS], pim[KS], psum[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int mi = 4 * ks + k4;
            const float2 pc = mi < L ? Pt[mi * 16 + n16] : make_float2(0.f, 0.f);
            pr[ks] = pc.x; pim[ks] = pc.y; psum[ks] = pc.x + pc.y;
        }
        const int so = n16 / LH, lh = n16 - (n16 / LH) * LH;
        const bool col_ok = so < SPN;
        floatx4 vr[2], vi[2];
#pragma unroll
        for (int bk = 0; bk < 2; ++bk) {
            const int s = (2 * wave + bk) * SPN + so;
            const float2* col = buf + (col_ok ? s * LH + lh : 0) * R2;
            floatx4 t1 = {0.f, 0.f, 0.f, 0.f}, t2 = {0.f, 0.f, 0.f, 0.f}, t3 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const float2 d = col[4 * ks + k4];
                t1 = __builtin_amdgcn_mfma_f32_16x16x4f32(pr[ks], d.x, t1, 0, 0, 0);
                t2 = __builtin_amdgcn_mfma_f32_16x16x4f32(pim[ks], d.y, t2, 0, 0, 0);
                t3 = __builtin_amdgcn_mfma_f32_16x16x4f32(psum[ks], d.x + d.y, t3, 0, 0, 0);
            }
            vr[bk] = t1 - t2;
            vi[bk] = t3 - t1 - t2;
        }
        __syncthreads();
        if (col_ok) {
#pragma unroll
            for (int bk = 0; bk < 2; ++bk) {
                const int s = (2 * wave + bk) * SPN + so;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    buf[(s * 16 + 4 * k4 + r) * Sh::R3 + lh] = make_float2(vr[bk][r], vi[bk][r]);
            }
        }
        __syncthreads();
    }

    // phase 4: right binning operator on the matrix cores, real part, scale, store.
    // B[Y][X] = sum_lx (Re V[Y][lx] Re Q'[X][lx] - Im V[Y][lx] Im Q'[X][lx]), Q' = Q~ with the
    // columns m_x > 0 counted twice, computed as B^T = Q' V^T: A = Q' (lane l: X = l % 16, lag
    // 4 ks + l / 16), B = V^T (lane l: row Y = l % 16, same lag), B^T[X][Y] out in lane
    // Y + 16 (X / 4), register X % 4: each lane holds 4 consecutive pixels of one CCD row.
    const int W = n * g.npx;
    double bsum = 0.0;
    {
        constexpr int KQ = (LH + 3) / 4;
        constexpr int SPW = TS / WAVES;               // subapertures per wave
        const float fscale = (float)scale;
        const bool vec4 = g.store_vec && (g.npx & 3) == 0;   // 16-byte aligned CCD rows
        static_assert(TS % WAVES == 0, "whole subapertures per wave");
        const int lane = t & 63, wave = t >> 6;
        const int n16 = lane & 15, k4 = lane >> 4;
        float qr[KQ], nqi[KQ];
#pragma unroll
        for (int ks = 0; ks < KQ; ++ks) {
            const int lh = 4 * ks + k4;
            const float2 qc = lh < LH ? Qc[n16 * L + NP - 1 + lh] : make_float2(0.f, 0.f);   // 16 rows
            const float w = lh == 0 ? 1.f : 2.f;
            qr[ks] = w * qc.x;
            nqi[ks] = -w * qc.y;
        }
        const int Y = n16, X0 = 4 * k4;
        // fused CoG: weights of the lane's pixels (x of its 4 columns, y of its row), its FP32
        // partials per subaperture (thresholded total, x and y moments), and, requested before the
        // binning products so that the loads overlap them, the slot and the sum(A^2) columns of
        // the subaperture whose sums this lane ends with (u_own, after the reduction below)
        float cxw[4] = {0.f, 0.f, 0.f, 0.f}, cyw = 0.f;
        float cpart[SPW][3];
        const int u_own = SPW == 4 ? (lane >> 4) : (lane >> 5), s_own = wave * SPW + u_own;
        int slot = -1;
        float sq = 0.f;
        if constexpr (COG != 0) {
            static_assert(SPW == 2 || SPW == 4, "2 or 4 subapertures per wave");
            if (Y < g.npx) {
                cyw = g.cogw[g.npx + Y];
#pragma unroll
                for (int r = 0; r < 4; ++r) cxw[r] = X0 + r < g.npx ? g.cogw[X0 + r] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < SPW; ++u) cpart[u][0] = cpart[u][1] = cpart[u][2] = 0.f;
            if (sub0 + s_own < g.n_sub) slot = g.slot[lenslet(s_own)];
            if (n16 < NP) sq = cog_sq[s_own * NP + n16];
        }
#pragma unroll
        for (int u = 0; u < SPW; ++u) {
            const int s = wave * SPW + u;
            if (sub0 + s >= g.n_sub) break;           // wave-uniform
            const int sub = lenslet(s);
            const float2* vrow = buf + (s * 16 + n16) * Sh::R3;
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < KQ; ++ks) {
                const int lh = 4 * ks + k4;
                const float2 v = lh < LH ? vrow[lh] : make_float2(0.f, 0.f);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(qr[ks], v.x, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(nqi[ks], v.y, acc, 0, 0, 0);
            }
            if (Y < g.npx && X0 < g.npx) {
                const int i = sub / n, j = sub - (sub / n) * n;
                float* dst = g.ccd + f * g.ccd_stride + (int64_t)(i * g.npx + Y) * W + (int64_t)j * g.npx + X0;
                if (!g.parseval) bsum += ((double)acc[0] + (double)acc[1]) + ((double)acc[2] + (double)acc[3]);
                const floatx4 o = acc * fscale;
                if (vec4) {
                    *reinterpret_cast<floatx4*>(dst) = o;
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (X0 + r < g.npx) dst[r] = o[r];
                }
                if constexpr (COG != 0) {
                    float w[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        float v = o[r];
                        if constexpr (COG == 1) v = clamp0(v - g.thr);          // subtract the threshold, clamp
                        else if constexpr (COG == 2) v = v < g.thr ? 0.f : v;  // pedestal
                        w[r] = X0 + r < g.npx ? v : 0.f;                        // columns beyond the subaperture
                    }
                    const float tot = (w[0] + w[1]) + (w[2] + w[3]);
                    cpart[u][0] = tot;
                    cpart[u][1] = fmaf(w[3], cxw[3], fmaf(w[2], cxw[2], fmaf(w[1], cxw[1], w[0] * cxw[0])));
                    cpart[u][2] = tot * cyw;
                }
            }
        }
        if constexpr (COG != 0) {
            // reduce each subaperture's partials over its 64 lanes (16 rows Y x 4 column groups):
            // the column groups in FP32 (lane bits 4-5, splitting the subapertures between the
            // halves and quarters of the wave so that one exchange serves two of them), then the
            // 16 rows within each row of lanes (DPP), the first step in FP32 (32 pixels per FP32
            // partial), the others in FP64.  Afterwards every lane of row r holds subaperture u_own.
            float c[3];
            const bool hi32 = lane >= 32, hi16 = (lane & 16) != 0;
            if constexpr (SPW == 2) {
#pragma unroll
                for (int v = 0; v < 3; ++v) {
                    const float keep = hi32 ? cpart[1][v] : cpart[0][v], send = hi32 ? cpart[0][v] : cpart[1][v];
                    c[v] = keep + __shfl_xor(send, 32, 64);
                    c[v] += __shfl_xor(c[v], 16, 64);
                }
            } else {
#pragma unroll
                for (int v = 0; v < 3; ++v) {
                    float a = hi32 ? cpart[2][v] : cpart[0][v], b = hi32 ? cpart[3][v] : cpart[1][v];
                    a += __shfl_xor(hi32 ? cpart[0][v] : cpart[2][v], 32, 64);
                    b += __shfl_xor(hi32 ? cpart[1][v] : cpart[3][v], 32, 64);
                    c[v] = (hi16 ? b : a) + __shfl_xor(hi16 ? a : b, 16, 64);
                }
            }
#pragma unroll
            for (int v = 0; v < 3; ++v) c[v] += dpp_f<kDppXor1>(c[v]);
            sq += dpp_f<kDppXor1>(sq);
            double m[4] = {c[0], c[1], c[2], sq};       // total, x and y moments, sum(A^2)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                m[v] += dpp_d<kDppXor2>(m[v]);
                m[v] += dpp_d<kDppHalfMirror>(m[v]);
                m[v] += dpp_d<kDppMirror>(m[v]);
            }
            if (n16 == 0 && (SPW == 4 || !hi16)) {       // one lane per subaperture
                const double m_f = m[3] * (scale * g.norm_const);   // Parseval flux of the subaperture
                if (slot >= 0) {
                    double* mo = g.mom + ((int64_t)f * g.n_valid + slot) * 4;
                    *reinterpret_cast<double2*>(mo) = make_double2(m_f, m[0]);
                    *reinterpret_cast<double2*>(mo + 2) = make_double2(m[1], m[2]);
                }
                // its share of the frame totals (flux, thresholded total)
                *reinterpret_cast<double2*>(g.tpart + ((int64_t)f * g.n_parts + tile * TS + s_own) * 2) =
                    slot >= 0 ? make_double2(m_f, m[0]) : make_double2(0.0, 0.0);
            }
        }
    }
    if (!g.parseval) {
        bsum = block_sum(bsum, red);
        if (t == 0) g.bpart[(int64_t)f * g.n_parts + tile] = bsum;
    }
}

template <int NP, int M, int TS, int COG>
__global__ void __launch_bounds__((LgsShape<NP, M, TS>::BT)) __attribute__((amdgpu_waves_per_eu(4, 8)))
sh_lgs_forward_kernel(NgsArgs g) {
    extern __shared__ float4 lds_raw[];
    __shared__ double red[4];
    lgs_tile<NP, M, TS, COG>(g, blockIdx.x, blockIdx.y, reinterpret_cast<float2*>(lds_raw), red);
}

// Lit flag of each subaperture (a nonzero pupil sample): one wave per subaperture, a ballot over
// its np x np samples.
__global__ void __launch_bounds__(256) subap_lit_kernel(const float* __restrict__ A, int ef, int n, int np, int n_sub,
                                                        uint8_t* __restrict__ lit) {
    const int sidx = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (sidx >= n_sub) return;                         // wave-uniform
    const int i = sidx / n, j = sidx - (sidx / n) * n;
    const float* a = A + (int64_t)(i * np) * ef + (int64_t)j * np;
    bool any = false;
    for (int e = lane; e < np * np; e += 64) any |= a[(int64_t)(e / np) * ef + (e % np)] != 0.f;
    const unsigned long long b = __ballot(any);
    if (lane == 0) lit[sidx] = b != 0 ? 1 : 0;
}

// Subaperture order of a static pupil: lit subapertures first, then the dark ones, each group in
// lenslet order (stable partition of the flags; one block, chunks of 1024).  Waves of the NGS
// kernels (tiles of the LGS kernel) then hold either lit or dark subapertures only, and dark ones
// skip the transforms.  Any permutation gives the same images: the kernels test darkness
// themselves.
__global__ void __launch_bounds__(1024) subap_order_kernel(const uint8_t* __restrict__ lit_flags, int n_sub,
                                                           int32_t* __restrict__ order) {
    __shared__ int wsum[16];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int n_lit = 0;
    for (int pass = 0; pass < 2; ++pass) {            // 0: count lit subapertures, 1: place them
        int lit_base = 0, dark_base = 0;
        for (int c0 = 0; c0 < n_sub; c0 += 1024) {
            const int sidx = c0 + t;
            const bool lit = sidx < n_sub && lit_flags[sidx] != 0;
            const unsigned long long b = __ballot(lit);
            const int in_wave = __popcll(b & ((1ull << lane) - 1ull));
            if (lane == 0) wsum[wv] = __popcll(b);
            __syncthreads();
            int before = 0, total = 0;
            for (int w = 0; w < 16; ++w) {
                const int c = wsum[w];
                before += (w < wv) ? c : 0;
                total += c;
            }
            __syncthreads();
            const int excl = before + in_wave;                         // lit ones ahead in the chunk
            if (pass == 1 && sidx < n_sub)
                order[lit ? lit_base + excl : n_lit + dark_base + (t - excl)] = sidx;
            lit_base += total;
            dark_base += min(1024, n_sub - c0) - total;
        }
        n_lit = lit_base;
    }
}

// Lit-first order of pupil A (ef x ef) for n x n subapertures of np x np samples; `lit` holds
// n_sub bytes of scratch.
int launch_subap_order(const float* A, int ef, int n, int np, int n_sub, uint8_t* lit, int32_t* order,
                       hipStream_t st) {
    hipLaunchKernelGGL(subap_lit_kernel, dim3((n_sub + 3) / 4), dim3(256), 0, st, A, ef, n, np, n_sub, lit);
    SHWFS_LAUNCH_CHECK();
    hipLaunchKernelGGL(subap_order_kernel, dim3(1), dim3(1024), 0, st, lit, n_sub, order);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

// Non-Parseval normalisation: scale = S0 * pixel_area * sum(A) / sum(B), then rescale in place.
__global__ void sh_rescale_kernel(float* __restrict__ ccd, int64_t ccd_stride, int64_t n_px,
                                  const double* __restrict__ bpart, int n_parts,
                                  const double* __restrict__ amp, int amp_fstep,
                                  double S0, double pixel_area) {
    __shared__ double red[16];
    const int f = blockIdx.y;
    double t = 0.0;
    for (int b = threadIdx.x; b < n_parts; b += blockDim.x) t += bpart[(int64_t)f * n_parts + b];
    t = block_sum(t, red);
    const double scale = S0 * pixel_area * amp[(int64_t)f * amp_fstep] / t;
    float* c = ccd + f * ccd_stride;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n_px;
         q += (int64_t)gridDim.x * blockDim.x)
        c[q] = (float)((double)c[q] * scale);
}

// ---------------------------------------------------------------- field resampling (fov_ovs != 1)
// SH.trigger_code, sh.py:350-360: the phase is first extended 1-2 pixels beyond the pupil edge
// (extrapolate_edge_pixel.py:4-29: pixel p with ext[p].x >= 0 takes v[ext.x], or
// 2 v[ext.x] - v[ext.y] when ext.y >= 0), then A and phase are resampled bilinearly on the
// ef_out grid (Interp2D default grid with rotation and shift, interp2d.py:48-122): output (Y, X)
// samples the input at (ay.x Y + ay.y X + ay.z, ax.x Y + ax.y X + ax.z), clamped to the input.
struct ResampleMap {
    double yy, yx, y0, xy, xx, x0;
};

__global__ void resample_kernel(const float* __restrict__ in, int64_t in_stride, const int2* __restrict__ ext,
                                float* __restrict__ out, int64_t out_stride, int ef_in, int ef_out,
                                ResampleMap m) {
    const int f = blockIdx.y;
    const float* v = in + f * in_stride;
    float* o = out + f * out_stride;
    const int64_t n_out = (int64_t)ef_out * ef_out;
    const double lim = ef_in - 1;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n_out; q += (int64_t)gridDim.x * blockDim.x) {
        const int Y = (int)(q / ef_out), X = (int)(q - (int64_t)Y * ef_out);
        const double y = fmin(fmax(m.yy * Y + m.yx * X + m.y0, 0.0), lim);
        const double x = fmin(fmax(m.xy * Y + m.xx * X + m.x0, 0.0), lim);
        const int i0 = min((int)floor(y), ef_in - 2), j0 = min((int)floor(x), ef_in - 2);
        const float ty = (float)(y - i0), tx = (float)(x - j0);
        float c[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int p = (i0 + (k >> 1)) * ef_in + j0 + (k & 1);
            float val = v[p];
            if (ext) {
                const int2 e = ext[p];
                if (e.x >= 0) val = e.y >= 0 ? 2.f * v[e.x] - v[e.y] : v[e.x];
            }
            c[k] = val;
        }
        o[q] = (1.f - ty) * ((1.f - tx) * c[0] + tx * c[1]) + ty * ((1.f - tx) * c[2] + tx * c[3]);
    }
}

// Transpose of resample_kernel (the resampling, edge extrapolation included, is linear in the
// phase): gin += R^T gout.  Each output pixel recomputes its taps exactly as the forward does and
// scatters its four bilinear weights times gout into the input pixels it read (2 w to the first
// and -w to the second source of an extrapolated pixel, at most 3 pixels away).  A block takes a
// tile of kRvjTY x kRvjTX output pixels; when the resampling magnifies (fov_ovs >= 1) their taps
// fall in a small input rectangle (the affine map takes its extremes at the tile corners, floor
// and clamp are monotonic), so the contributions accumulate in LDS and each touched input pixel
// receives one global atomic per tile instead of one per contribution.  A rectangle larger than
// the LDS budget, and any contribution outside it, goes to global atomics directly.  Float
// atomics: the summation order is not fixed, so results agree with a fixed-order sum to FP32
// rounding.  gin must be zeroed.  Grid (tiles, frames), 256 threads.
constexpr int kRvjTY = 16, kRvjTX = 128, kRvjLds = 8192;
__global__ void __launch_bounds__(256)
resample_vjp_kernel(const float* __restrict__ gout, int64_t gout_stride, const int2* __restrict__ ext,
                    float* __restrict__ gin, int64_t gin_stride, int ef_in, int ef_out, ResampleMap m, int tiles_x) {
    __shared__ float acc[kRvjLds];
    const int f = blockIdx.y;
    const float* go = gout + f * gout_stride;
    float* gi = gin + f * gin_stride;
    const int ty = blockIdx.x / tiles_x, tx = blockIdx.x - (blockIdx.x / tiles_x) * tiles_x;
    const int Y0 = ty * kRvjTY, X0 = tx * kRvjTX;
    const int Y1 = min(Y0 + kRvjTY, ef_out) - 1, X1 = min(X0 + kRvjTX, ef_out) - 1;
    const double lim = ef_in - 1;
    double ylo = lim, yhi = 0.0, xlo = lim, xhi = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int Yc = (k & 1) ? Y1 : Y0, Xc = (k & 2) ? X1 : X0;
        const double y = fmin(fmax(m.yy * Yc + m.yx * Xc + m.y0, 0.0), lim);
        const double x = fmin(fmax(m.xy * Yc + m.xx * Xc + m.x0, 0.0), lim);
        ylo = fmin(ylo, y); yhi = fmax(yhi, y);
        xlo = fmin(xlo, x); xhi = fmax(xhi, x);
    }
    const int r0 = max(min((int)floor(ylo), ef_in - 2) - 3, 0);
    const int r1 = min(min((int)floor(yhi), ef_in - 2) + 4, ef_in - 1);
    const int c0 = max(min((int)floor(xlo), ef_in - 2) - 3, 0);
    const int c1 = min(min((int)floor(xhi), ef_in - 2) + 4, ef_in - 1);
    const int fw = c1 - c0 + 1;
    const int fn = (r1 - r0 + 1) * fw;
    const bool local = fn <= kRvjLds;                                  // block-uniform
    if (local)
        for (int e = threadIdx.x; e < fn; e += blockDim.x) acc[e] = 0.f;
    __syncthreads();
    auto add = [&](int r, int c, float v) {
        if (local && r >= r0 && r <= r1 && c >= c0 && c <= c1) atomicAdd(&acc[(r - r0) * fw + (c - c0)], v);
        else atomicAdd(gi + (int64_t)r * ef_in + c, v);
    };
    for (int e = threadIdx.x; e < kRvjTY * kRvjTX; e += blockDim.x) {
        const int Y = Y0 + e / kRvjTX, X = X0 + (e - (e / kRvjTX) * kRvjTX);
        if (Y > Y1 || X > X1) continue;
        const float gq = go[(int64_t)Y * ef_out + X];
        if (gq == 0.f) continue;
        const double y = fmin(fmax(m.yy * Y + m.yx * X + m.y0, 0.0), lim);
        const double x = fmin(fmax(m.xy * Y + m.xx * X + m.x0, 0.0), lim);
        const int i0 = min((int)floor(y), ef_in - 2), j0 = min((int)floor(x), ef_in - 2);
        const float ty_ = (float)(y - i0), tx_ = (float)(x - j0);
        const float w[4] = {(1.f - ty_) * (1.f - tx_), (1.f - ty_) * tx_, ty_ * (1.f - tx_), ty_ * tx_};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float wg = w[k] * gq;
            if (wg == 0.f) continue;
            const int r = i0 + (k >> 1), c = j0 + (k & 1);
            const int2 src = ext ? ext[r * ef_in + c] : make_int2(-1, -1);
            if (src.x < 0) {
                add(r, c, wg);
            } else if (src.y < 0) {
                add(src.x / ef_in, src.x - (src.x / ef_in) * ef_in, wg);
            } else {
                add(src.x / ef_in, src.x - (src.x / ef_in) * ef_in, 2.f * wg);
                add(src.y / ef_in, src.y - (src.y / ef_in) * ef_in, -wg);
            }
        }
    }
    if (!local) return;                                                 // block-uniform
    __syncthreads();
    for (int e = threadIdx.x; e < fn; e += blockDim.x) {
        const float v = acc[e];
        if (v != 0.f) atomicAdd(gi + (int64_t)(r0 + e / fw) * ef_in + c0 + (e - (e / fw) * fw), v);
    }
}

// Zero n_px floats of each frame (frame stride `stride`); grid (blocks, frames).
__global__ void zero_frames_kernel(float* __restrict__ p, int64_t stride, int64_t n_px) {
    float* q = p + blockIdx.y * stride;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_px; i += (int64_t)gridDim.x * blockDim.x)
        q[i] = 0.f;
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


// Fused-path variant: the frame totals come as per-wave partial sums from the SH kernel, so every
// block reduces them (same fixed order in every block) and handles its own slice of subapertures
// (grid-stride over them).  grid = (a few blocks, n_frames).
// Slopes, flux and totals of frame f from the fused-CoG moments (sh_slopec.py:365-438): the
// frame totals from the n_parts partial pairs, then per subaperture factor = 1 / total, set to 0
// below 1e-3 of the mean total (sh_slopec.py:393-400).  One block row per frame.
__device__ __forceinline__ void cog_finalize_frame(const double* __restrict__ mom, const double* __restrict__ tpart,
                                                   int n_parts, int n_sub, float* __restrict__ slopes,
                                                   int64_t slopes_stride, double* __restrict__ flux,
                                                   double* __restrict__ totals, double* __restrict__ factor_out,
                                                   int f, double* red) {
    const double* tp = tpart + (int64_t)f * n_parts * 2;
    double sf = 0.0, st = 0.0;
    for (int b = threadIdx.x; b < n_parts; b += blockDim.x) {
        const double2 v = *reinterpret_cast<const double2*>(tp + 2 * b);
        sf += v.x;
        st += v.y;
    }
    const double2 sft = block_sum2(sf, st, red);
    sf = sft.x;
    st = sft.y;
    const double limit = 1.0 / ((st / n_sub) * 1e-3);
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n_sub; s += gridDim.x * blockDim.x) {
        const double* m = mom + ((int64_t)f * n_sub + s) * 4;
        const double2 m01 = *reinterpret_cast<const double2*>(m);
        const double2 m23 = *reinterpret_cast<const double2*>(m + 2);
        double factor = 1.0 / m01.y;
        if (factor > limit) factor = 0.0;
        if (factor_out) factor_out[(int64_t)f * n_sub + s] = factor;
        float* sl = slopes + f * slopes_stride;
        sl[s] = (float)(m23.x * factor);
        sl[n_sub + s] = (float)(m23.y * factor);
        flux[(int64_t)f * n_sub + s] = m01.x;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        totals[2 * f] = sf;
        totals[2 * f + 1] = sf / n_sub;
    }
}

__global__ void cog_finalize_parts_kernel(const double* __restrict__ mom, const double* __restrict__ tpart,
                                          int n_parts, int n_sub, float* __restrict__ slopes,
                                          int64_t slopes_stride, double* __restrict__ flux,
                                          double* __restrict__ totals, double* __restrict__ factor_out) {
    __shared__ double red[32];
    cog_finalize_frame(mom, tpart, n_parts, n_sub, slopes, slopes_stride, flux, totals, factor_out, blockIdx.y, red);
}

// The same for every fused-CoG sensor of a multi-sensor batch in one launch: block row y serves
// frame map[y].y of record map[y].x.
struct CogFinalizeRec {
    const double* mom;
    const double* tpart;
    int n_parts, n_sub;
    float* slopes;
    int64_t slopes_stride;
    double* flux;
    double* totals;
};

__global__ void cog_finalize_batch_kernel(const CogFinalizeRec* __restrict__ recs, const int2* __restrict__ map,
                                          int n_rows) {
    __shared__ double red[32];
    for (int y = blockIdx.y; y < n_rows; y += gridDim.y) {
        const int2 rf = map[y];
        const CogFinalizeRec& r = recs[rf.x];
        if ((int)blockIdx.x * (int)blockDim.x < r.n_sub || blockIdx.x == 0)
            cog_finalize_frame(r.mom, r.tpart, r.n_parts, r.n_sub, r.slopes, r.slopes_stride, r.flux, r.totals,
                               nullptr, rf.y, red);
        __syncthreads();                             // red is reused by the next row
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
    const double2* cg;     // non-Parseval plans: per frame (scale c_f, cotangent offset alpha_f)
};

// Per-frame factors of the VJP of an image normalised by its own sum (non-Parseval plans): with
// I = c B, c = S0 pixel_area sum(A) / sum(B), the image VJP is the fixed-scale VJP of the
// cotangent g' = g - alpha, alpha = <g, I> / sum(I) (the derivative of c(phi) folds into the
// offset).  One block per frame; `img` is the forward image I of the same inputs, `bpart` the
// forward's partial sums of the unscaled B.
__global__ void vjp_scale_kernel(const double* __restrict__ bpart, int n_parts, const double* __restrict__ amp,
                                 int amp_fstep, double S0, double pixel_area, const float* __restrict__ img,
                                 int64_t img_stride, const float* __restrict__ gccd, int64_t gccd_stride,
                                 int64_t n_px, double2* __restrict__ cg) {
    __shared__ double red[16];
    const int f = blockIdx.x;
    double t = 0.0;
    for (int b = threadIdx.x; b < n_parts; b += blockDim.x) t += bpart[(int64_t)f * n_parts + b];
    const float* im = img + f * img_stride;
    const float* gc = gccd + f * gccd_stride;
    double gi = 0.0, si = 0.0;
    for (int64_t q = threadIdx.x; q < n_px; q += blockDim.x) {
        const double v = im[q];
        gi += (double)gc[q] * v;
        si += v;
    }
    t = block_sum(t, red);
    const double2 gs = block_sum2(gi, si, red);
    if (threadIdx.x == 0) cg[f] = make_double2(S0 * pixel_area * amp[(int64_t)f * amp_fstep] / t, gs.x / gs.y);
}

#include "sh_general_vjp.h"

// SH backward for the NGS path with Parseval normalisation (I = c B, c independent of the
// phase): with G_B = c dL/dI on the subaperture tile,
//   G_S = P^T G_B Q,   X = G_S conj(F),   H = FFT_M(X) on the np x np support,
//   dL/dphi = -2 (2 pi / lambda) Im(H E).
// NPX > 0: CCD pixels per subaperture side fixed at compile time (loops unrolled); 0: g.npx.
// The binning operators P, Q are staged in LDS after the G_B tiles.
template <int NP, int M, int TS, int NPX>
__global__ void __launch_bounds__(TS * NP)
sh_ngs_vjp_kernel(VjpArgs g) {
    extern __shared__ float4 lds_raw[];
    constexpr int G1S = M * NP + 1;
    constexpr int NT = TS * NP;
    constexpr int CPT = (TS * M + NT - 1) / NT;
    const int npx = NPX > 0 ? NPX : g.npx;
    float2* buf = reinterpret_cast<float2*>(lds_raw);
    float* gb = reinterpret_cast<float*>(buf + TS * G1S);
    const int npx2 = npx * npx;
    float* Pl = gb + TS * npx2;                      // P: npx x M
    float* Ql = Pl + npx * M;                        // Q: npx x M
    for (int e = threadIdx.x; e < npx * M; e += NT) {
        Pl[e] = g.roww[e];
        Ql[e] = g.colw[e];
    }
    const int t = threadIdx.x, f = blockIdx.y, tile = blockIdx.x, n = g.n_lenses;
    const double* am = g.amp + (int64_t)f * g.amp_fstep;
    float c, alpha = 0.f;
    if (g.cg) {
        c = (float)g.cg[f].x;
        alpha = (float)g.cg[f].y;
    } else {
        c = (float)(g.S0 * g.pixel_area * am[0] / (g.norm_const * am[1]));
    }

    for (int e = t; e < TS * npx2; e += NT) {         // G_B tiles
        const int s = e / npx2, rem = e - s * npx2;
        const int Y = rem / npx, X = rem - Y * npx;
        const int sub = tile * TS + s;
        float v = 0.f;
        if (sub < g.n_sub) {
            const int i = sub / n, j = sub - (sub / n) * n;
            v = (g.gccd[f * g.gccd_stride + (int64_t)(i * npx + Y) * (n * npx) + j * npx + X] - alpha) * c;
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
            if (NPX > 0) {
                float qc[NPX > 0 ? NPX : 1];
#pragma unroll
                for (int X = 0; X < NPX; ++X) qc[X] = Ql[X * M + kx];
#pragma unroll
                for (int Y = 0; Y < NPX; ++Y) {
                    float u = 0.f;
#pragma unroll
                    for (int X = 0; X < NPX; ++X) u = fmaf(gbs[Y * NPX + X], qc[X], u);
#pragma unroll
                    for (int ky = 0; ky < M; ++ky) gs[ky] = fmaf(Pl[Y * M + ky], u, gs[ky]);
                }
            } else {
                for (int Y = 0; Y < npx; ++Y) {
                    float u = 0.f;
                    for (int X = 0; X < npx; ++X) u = fmaf(gbs[Y * npx + X], Ql[X * M + kx], u);
#pragma unroll
                    for (int ky = 0; ky < M; ++ky) gs[ky] = fmaf(Pl[Y * M + ky], u, gs[ky]);
                }
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
// Block = one strip of NP pupil rows x a chunk of modes; one thread per pupil column.  Only the
// columns of valid subapertures read the modes (the Jacobian is zero elsewhere), and MU modes are
// loaded before they are reduced so that each wave keeps several row segments in flight.
template <int NP>
__global__ void im_from_jacobian_kernel(const float* __restrict__ jx, const float* __restrict__ jy, int ef,
                                        int n_lenses, const float* __restrict__ modes, int64_t mode_stride,
                                        int n_modes, int modes_per_block, const int32_t* __restrict__ slot,
                                        int n_valid, float scale, float* __restrict__ im) {
    constexpr int MU = 4;
    const int i = blockIdx.x;
    const int col = threadIdx.x;
    const int j = col / NP, xin = col - (col / NP) * NP;
    const int sl = (col < ef) ? slot[i * n_lenses + j] : -1;
    const bool valid = sl >= 0 && sl < n_valid;      // slots outside [0, n_valid) are ignored
    float jxr[NP], jyr[NP];
#pragma unroll
    for (int r = 0; r < NP; ++r) {
        const int64_t o = (int64_t)(i * NP + r) * ef + col;
        jxr[r] = valid ? jx[o] : 0.f;
        jyr[r] = valid ? jy[o] : 0.f;
    }
    const int m0 = blockIdx.y * modes_per_block;
    const int m1 = min(n_modes, m0 + modes_per_block);
    const float* base = modes + (int64_t)(i * NP) * ef + col;
    for (int mb = m0; mb < m1; mb += MU) {
        float v[MU][NP];
#pragma unroll
        for (int u = 0; u < MU; ++u) {
            const bool ld = valid && mb + u < m1;
            const float* md = base + (int64_t)(mb + u) * mode_stride;
#pragma unroll
            for (int r = 0; r < NP; ++r) v[u][r] = ld ? md[(int64_t)r * ef] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < MU; ++u) {
            float px = 0.f, py = 0.f;
#pragma unroll
            for (int r = 0; r < NP; ++r) {
                px = fmaf(jxr[r], v[u][r], px);
                py = fmaf(jyr[r], v[u][r], py);
            }
#pragma unroll
            for (int off = NP / 2; off > 0; off >>= 1) {
                px += __shfl_xor(px, off, kWave);
                py += __shfl_xor(py, off, kWave);
            }
            if (valid && xin == 0 && mb + u < m1) {
                im[(int64_t)(mb + u) * 2 * n_valid + sl] = scale * px;
                im[(int64_t)(mb + u) * 2 * n_valid + n_valid + sl] = scale * py;
            }
        }
    }
}


// Same contraction with float4 traffic: each lane holds a 4 x 4 pixel block (4 columns of 4
// rows) of a subaperture, so a subaperture spans (NP/4)^2 consecutive lanes; MU modes in flight.
// GRID: results go to grid[m][2][n_lenses^2] in subaperture-grid order (consecutive lanes write
// consecutive words) instead of to the slot-ordered im rows; im_gather_kernel reorders them.
template <int NP, bool GRID>
__global__ void __launch_bounds__(1024)
im_from_jacobian_v4_kernel(const float* __restrict__ jx, const float* __restrict__ jy, int ef, int n_lenses,
                           const float* __restrict__ modes, int64_t mode_stride, int n_modes, int modes_per_block,
                           const int32_t* __restrict__ slot, int n_valid, float scale, float* __restrict__ im) {
    constexpr int MU = 2;
    constexpr int RS = NP / 4;                     // row blocks per subaperture
    constexpr int LPS = RS * (NP / 4);             // lanes per subaperture
    const int i = blockIdx.x;
    const int t = threadIdx.x;
    const int rs = t % RS, qd = t / RS;
    const int col = qd * 4;
    const int sl = (col < ef) ? slot[i * n_lenses + col / NP] : -1;
    const bool valid = sl >= 0 && sl < n_valid;      // slots outside [0, n_valid) are ignored
    const int64_t row0 = (int64_t)(i * NP + rs * 4) * ef + col;
    float4 jxr[4], jyr[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int64_t o = row0 + (int64_t)r * ef;
        jxr[r] = valid ? *reinterpret_cast<const float4*>(jx + o) : make_float4(0.f, 0.f, 0.f, 0.f);
        jyr[r] = valid ? *reinterpret_cast<const float4*>(jy + o) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const int m0 = blockIdx.y * modes_per_block;
    const int m1 = min(n_modes, m0 + modes_per_block);
    const float* base = modes + row0;
    for (int mb = m0; mb < m1; mb += MU) {
        float4 v[MU][4];
#pragma unroll
        for (int u = 0; u < MU; ++u) {
            const bool ld = valid && mb + u < m1;
            const float* md = base + (int64_t)(mb + u) * mode_stride;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (ld) {
                    const floatx4 w = __builtin_nontemporal_load(reinterpret_cast<const floatx4*>(md + (int64_t)r * ef));
                    v[u][r] = make_float4(w.x, w.y, w.z, w.w);
                } else {
                    v[u][r] = make_float4(0.f, 0.f, 0.f, 0.f);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < MU; ++u) {
            float px = 0.f, py = 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                px = fmaf(jxr[r].x, v[u][r].x, fmaf(jxr[r].y, v[u][r].y, fmaf(jxr[r].z, v[u][r].z, fmaf(jxr[r].w, v[u][r].w, px))));
                py = fmaf(jyr[r].x, v[u][r].x, fmaf(jyr[r].y, v[u][r].y, fmaf(jyr[r].z, v[u][r].z, fmaf(jyr[r].w, v[u][r].w, py))));
            }
#pragma unroll
            for (int off = LPS / 2; off > 0; off >>= 1) {
                px += __shfl_xor(px, off, kWave);
                py += __shfl_xor(py, off, kWave);
            }
            if (valid && (t % LPS) == 0 && mb + u < m1) {
                if (GRID) {
                    const int64_t nn = (int64_t)n_lenses * n_lenses;
                    const int64_t gidx = (int64_t)i * n_lenses + col / NP;
                    im[(int64_t)(mb + u) * 2 * nn + gidx] = scale * px;
                    im[(int64_t)(mb + u) * 2 * nn + nn + gidx] = scale * py;
                } else {
                    im[(int64_t)(mb + u) * 2 * n_valid + sl] = scale * px;
                    im[(int64_t)(mb + u) * 2 * n_valid + n_valid + sl] = scale * py;
                }
            }
        }
    }
}

// inv[slot[g]] = g for every subaperture g of the grid whose slot lies in [0, n_valid); slots
// that no subaperture claims keep -1 (their rows come out zero).  One block: the -1 fill is
// ordered before the scatter by the block barrier.
__global__ void __launch_bounds__(1024) slot_inverse_kernel(const int32_t* __restrict__ slot, int nn, int n_valid,
                                                            int32_t* __restrict__ inv) {
    for (int s = threadIdx.x; s < n_valid; s += blockDim.x) inv[s] = -1;
    __syncthreads();
    for (int g = threadIdx.x; g < nn; g += blockDim.x) {
        const int s = slot[g];
        if (s >= 0 && s < n_valid) inv[s] = g;
    }
}

// im[m][s] = grid[m][0][inv[s]], im[m][n_valid + s] = grid[m][1][inv[s]].  One block per mode:
// the mode's grid row is read into LDS with coalesced loads, then written in slot order with
// coalesced stores (the permutation happens in LDS).  Needs 2 nn floats of LDS.
__global__ void __launch_bounds__(256)
im_gather_lds_kernel(const float* __restrict__ grid, const int32_t* __restrict__ inv, int nn, int n_valid,
                     float* __restrict__ im) {
    extern __shared__ float4 lds_raw[];
    float* row_l = reinterpret_cast<float*>(lds_raw);
    const int m = blockIdx.x;
    const float* gx = grid + (int64_t)m * 2 * nn;
    if ((nn & 1) == 0) {                               // rows 16-byte aligned: float4 fill, 4 in flight
        const int n4 = nn >> 1;                        // 2 nn floats = nn / 2 float4
        const float4* g4 = reinterpret_cast<const float4*>(gx);
        float4* l4 = reinterpret_cast<float4*>(row_l);
        for (int e = threadIdx.x; e < n4; e += 4 * blockDim.x) {
            float4 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = e + u * (int)blockDim.x;
                v[u] = k < n4 ? g4[k] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = e + u * (int)blockDim.x;
                if (k < n4) l4[k] = v[u];
            }
        }
    } else {
        for (int e = threadIdx.x; e < 2 * nn; e += blockDim.x) row_l[e] = gx[e];
    }
    __syncthreads();
    float* row = im + (int64_t)m * 2 * n_valid;
    for (int s = threadIdx.x; s < n_valid; s += blockDim.x) {
        const int g = inv[s];
        row[s] = g >= 0 ? row_l[g] : 0.f;
        row[n_valid + s] = g >= 0 ? row_l[nn + g] : 0.f;
    }
}

// Same contraction for any np_sub and pupil width (the any-size path's sensors): one block per
// (subaperture, chunk of modes); its threads stride over the subaperture's np x np pixels, 8 modes
// at a time, and the block reduces the 16 partial sums.  Subapertures without a valid slot exit.
__global__ void __launch_bounds__(256)
im_from_jacobian_any_kernel(const float* __restrict__ jx, const float* __restrict__ jy, int ef, int n_lenses, int np,
                            const float* __restrict__ modes, int64_t mode_stride, int n_modes, int modes_per_block,
                            const int32_t* __restrict__ slot, int n_valid, float scale, float* __restrict__ im) {
    constexpr int MU = 8, NW = 256 / kWave;
    __shared__ float red[NW][2 * MU];
    const int sub = blockIdx.x;
    const int sl = slot[sub];
    if (sl < 0 || sl >= n_valid) return;                 // the whole block
    const int i = sub / n_lenses, j = sub - (sub / n_lenses) * n_lenses;
    const int64_t off = (int64_t)(i * np) * ef + (int64_t)j * np;
    const int npp = np * np, w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int m0 = blockIdx.y * modes_per_block;
    const int m1 = min(n_modes, m0 + modes_per_block);
    for (int mb = m0; mb < m1; mb += MU) {
        float px[MU], py[MU];
#pragma unroll
        for (int u = 0; u < MU; ++u) px[u] = py[u] = 0.f;
        for (int e = threadIdx.x; e < npp; e += blockDim.x) {
            const int y = e / np, x = e - (e / np) * np;
            const int64_t o = off + (int64_t)y * ef + x;
            const float a = jx[o], b = jy[o];
#pragma unroll
            for (int u = 0; u < MU; ++u) {
                const float v = mb + u < m1 ? modes[(int64_t)(mb + u) * mode_stride + o] : 0.f;
                px[u] = fmaf(a, v, px[u]);
                py[u] = fmaf(b, v, py[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < MU; ++u)
            for (int d = kWave / 2; d > 0; d >>= 1) {
                px[u] += __shfl_xor(px[u], d, kWave);
                py[u] += __shfl_xor(py[u], d, kWave);
            }
        if (lane == 0) {
#pragma unroll
            for (int u = 0; u < MU; ++u) {
                red[w][u] = px[u];
                red[w][MU + u] = py[u];
            }
        }
        __syncthreads();
        if (threadIdx.x < 2 * MU) {
            const int u = threadIdx.x % MU, c = threadIdx.x / MU;
            float t = 0.f;
            for (int q = 0; q < NW; ++q) t += red[q][threadIdx.x];
            if (mb + u < m1) im[(int64_t)(mb + u) * 2 * n_valid + c * n_valid + sl] = scale * t;
        }
        __syncthreads();
    }
}

// Same permutation without LDS (grids too large for it): one thread per slot and mode.
__global__ void im_gather_kernel(const float* __restrict__ grid, const int32_t* __restrict__ inv, int nn, int n_valid,
                                 float* __restrict__ im) {
    const int m = blockIdx.y;
    const float* gx = grid + (int64_t)m * 2 * nn;
    float* row = im + (int64_t)m * 2 * n_valid;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n_valid; s += gridDim.x * blockDim.x) {
        const int g = inv[s];
        row[s] = g >= 0 ? gx[g] : 0.f;
        row[n_valid + s] = g >= 0 ? gx[nn + g] : 0.f;
    }
}


// ---------------------------------------------------------------- multi-sensor persistent batch
// One launch processes a work list of (sensor, frame, tile) items of uneven shapes: NGS tiles
// (np 8, MFMA binning, 32 subapertures) and LGS tiles (np 16, N 128, 8 subapertures).  Blocks
// stride over the list; LDS is sized for the larger body.  Per-sensor completion times are
// recorded with the constant-rate wall clock for latency statistics.
struct BatchItem {
    int32_t sensor, frame, tile, nf;     // frames [frame, frame + nf) of one subaperture tile
};

constexpr int kBatchNgsLds = NgsMfmaShape<8>::LDS_F4 * 16;
constexpr int kBatchLgsLds = LgsShape<16, 32, 8>::LDS_F2 * 8 + LgsShape<16, 32, 8>::COG_BYTES;
constexpr int kBatchNgsFpb = 4;          // frames streamed per NGS item (next-frame prefetch)

__global__ void batch_stamp_kernel(unsigned long long* __restrict__ ts, int n_sensors) {
    if (threadIdx.x == 0) ts[0] = (unsigned long long)wall_clock64();
    for (int s = threadIdx.x; s < n_sensors; s += blockDim.x) ts[1 + s] = 0ull;
}

// Completion stamp of a sensor that the batch runs with its plan's own launch.
__global__ void batch_sensor_stamp_kernel(unsigned long long* __restrict__ ts, int sensor) {
    if (threadIdx.x == 0) ts[1 + sensor] = (unsigned long long)wall_clock64();
}

// Body variant of a batch sensor: which instantiation of its body (KIND 0: NGS MFMA np 8 tiles,
// KIND 1: LGS np 16 tiles) its items run.  NGS: 0 image normalised by its own sum, 1-3 fused CoG
// with 16 px per subaperture (no threshold, subtracted threshold, pedestal), 4-6 the same with 8 px,
// 7 / 8 plain with 16 / 8 px and 16-byte aligned CCD rows, 9 any other; LGS: 0 plain, 1-3 fused
// CoG (no threshold, subtracted threshold, pedestal).
constexpr int kBatchVariants[2] = {10, 4};
__host__ __device__ inline int batch_variant(const NgsArgs& g, int kind) {
    const int mode = (g.cog == 1 && g.thr == 0.f) ? 3 : g.cog;
    if (kind == 1) return g.cog ? (mode == 3 ? 1 : mode + 1) : 0;
    if (!g.parseval) return 0;
    if (g.cog) return (g.npx == 16 ? 1 : 4) + (mode == 3 ? 0 : mode);
    if (g.npx == 16 && g.store_vec) return 7;
    if (g.npx == 8 && g.store_vec) return 8;
    return 9;
}

template <int VAR>
__device__ __forceinline__ void batch_ngs_body(const NgsArgs& g, int wave_id, int f0, int f1, float* sw) {
    if constexpr (VAR == 0) ngs_mfma_frames<8, false, 0, 0, false>(g, wave_id, f0, f1, sw);
    else if constexpr (VAR == 1) ngs_mfma_frames<8, true, 16, 3, false>(g, wave_id, f0, f1, sw);
    else if constexpr (VAR == 2) ngs_mfma_frames<8, true, 16, 1, false>(g, wave_id, f0, f1, sw);
    else if constexpr (VAR == 3) ngs_mfma_frames<8, true, 16, 2, false>(g, wave_id, f0, f1, sw);
    else if constexpr (VAR == 4) ngs_mfma_frames<8, true, 8, 3, false>(g, wave_id, f0, f1, sw);
    else if constexpr (VAR == 5) ngs_mfma_frames<8, true, 8, 1, false>(g, wave_id, f0, f1, sw);
    else if constexpr (VAR == 6) ngs_mfma_frames<8, true, 8, 2, false>(g, wave_id, f0, f1, sw);
    else if constexpr (VAR == 7) ngs_mfma_frames<8, true, 16, 0, false>(g, wave_id, f0, f1, sw);
    else if constexpr (VAR == 8) ngs_mfma_frames<8, true, 8, 0, false>(g, wave_id, f0, f1, sw);
    else ngs_mfma_frames<8, true, 0, 0, false>(g, wave_id, f0, f1, sw);
}

// One persistent launch per body type and variant present, so that each runs with its own
// register and LDS footprint (one kernel holding several variants spills registers); blocks
// stride over the variant's items (grouped by shwfs_batch's layout).
template <int KIND, int VAR>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
sh_batch_kernel(const NgsArgs* __restrict__ sensors, const BatchItem* __restrict__ items, int n_items,
                unsigned long long* __restrict__ ts) {
    extern __shared__ float4 lds_raw[];
    __shared__ double red[4];
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        __syncthreads();                               // LDS of the previous item is free
        const BatchItem w = items[it];
        const NgsArgs& g = sensors[w.sensor];
        if constexpr (KIND == 0) {
            const int wv = threadIdx.x >> 6;
            float* sw = reinterpret_cast<float*>(lds_raw + wv * NgsMfmaShape<8>::WAVE_F4);
            batch_ngs_body<VAR>(g, w.tile * NgsMfmaShape<8>::WPB + wv, w.frame, w.frame + w.nf, sw);
        } else {
            constexpr int COG = VAR == 0 ? 0 : VAR == 1 ? 3 : VAR - 1;
            lgs_tile<16, 32, 8, COG>(g, w.tile, w.frame, reinterpret_cast<float2*>(lds_raw), red);
        }
        __syncthreads();
        if (threadIdx.x == 0) atomicMax(&ts[1 + w.sensor], (unsigned long long)wall_clock64());
    }
}

using BatchKernel = void (*)(const NgsArgs*, const BatchItem*, int, unsigned long long*);

template <int KIND, int... V>
BatchKernel batch_kernel_of(int var, std::integer_sequence<int, V...>) {
    BatchKernel k = nullptr;
    ((k = (var == V) ? sh_batch_kernel<KIND, V> : k), ...);
    return k;
}

}  // namespace

// ---------------------------------------------------------------- push-pull calibration fields
// Fields of a modal push-pull (im_rec_calibrator.py:58-63 driven by modal_pushpull_signal.py):
// out[k] = phi0 + a m_k (push), out[K + k] = phi0 - a m_k (pull); grid (blocks, K).
__global__ void pushpull_fields_kernel(const float* __restrict__ phi0, const float* __restrict__ modes,
                                       int64_t mode_stride, int64_t n_px, float amp, float* __restrict__ out,
                                       int64_t out_stride, int K) {
    const int k = blockIdx.y;
    const float* m = modes + k * mode_stride;
    float* op = out + k * out_stride;
    float* om = out + (k + (int64_t)K) * out_stride;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n_px; q += (int64_t)gridDim.x * blockDim.x) {
        const float p = phi0[q], d = amp * m[q];
        op[q] = p + d;
        om[q] = p - d;
    }
}

// Interaction-matrix rows from the push-pull slopes as ImRecCalibrator accumulates them
// (im_rec_calibrator.py:58-63: im[mode] += slopes / command, for +a then -a):
// im[k] = s[k] / a + s[K + k] / (-a).
__global__ void pushpull_rows_kernel(const float* __restrict__ slopes, int64_t s_stride, int n, float amp,
                                     float* __restrict__ im, int64_t im_stride, int K) {
    const int k = blockIdx.y;
    const float* sp = slopes + k * s_stride;
    const float* sm = slopes + (k + (int64_t)K) * s_stride;
    float* r = im + k * im_stride;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
        r[q] = sp[q] / amp + sm[q] / (-amp);
}

// ============================================================================ plans
struct shwfs_plan {
    shwfs_geom g{};
    int device = 0;
    int n_sub = 0;
    int npx_pad = 0;
    int ts = 0;
    int n_tiles = 0;
    int n_parts = 0;               // non-Parseval partial sums per frame
    int n_cog_parts = 0;           // fused-CoG frame-total partials per frame (LGS: one per tile position)
    int amp_blocks = 0;
    int rows_alloc = 0;            // rows of the uploaded binning operators (zero padded)
    bool mfma = false;             // binning on the matrix cores (sh_ngs_mfma_kernel)
    bool general = false;          // any-size path (sh_general_kernel, variant 4)
    int terms = 1;                 // separable terms of the FoV mask (> 1: general path only)
    float2* d_tw = nullptr;        // general path: [M/2] twiddles
    float* d_slab = nullptr;       // general path: [gen_blocks][slab_floats] scratch
    int gen_blocks = 0;
    int64_t slab_floats = 0;
    bool bf3 = false;              // MFMA binning from bf16 pieces (three-term split, variant 3)
    bool kernels_set = false;      // LGS: kernel spectra uploaded
    int fpb = 4;                   // frames per block of the streaming MFMA kernel (1: no streaming)
    int L = 0;                     // LGS lag-grid size 2 np - 1
    float* d_roww = nullptr;
    float* d_rowlag = nullptr;     // MFMA path: row operator in the lag domain, P~ = P A (16 x 16)
    float* d_colw = nullptr;
    float* d_kspec = nullptr;
    double* d_amp = nullptr;       // [max_frames][2] statistics of the last unregistered amplitude
    float* d_scale = nullptr;      // [max_frames] its Parseval flux scale
    double* d_amp_st = nullptr;    // [max_frames][2] statistics of the registered (static) pupil
    float* d_scale_st = nullptr;   // [max_frames] its Parseval flux scale
    int ef_in = 0;                 // resampling: side of the input field (0: no resampling)
    ResampleMap rmap{};
    int2* d_ext = nullptr;         // [ef_in^2] edge-extrapolation sources of the phase
    float* d_A_rs = nullptr;       // [max_frames][ef_size^2] resampled amplitude
    float* d_A_rs_st = nullptr;    // [static_frames][ef_size^2] resampled registered pupil
    int A_rs_st_frames = 0;        // frames allocated in d_A_rs_st
    float* d_ph_rs = nullptr;      // [max_frames][ef_size^2] resampled phase
    float* d_gph_rs = nullptr;     // VJP of a resampled plan: [max_frames][ef_size^2] sensor-grid gradient
    float* d_vjp_ccd = nullptr;    // VJP of a non-Parseval plan: [max_frames][ccd px] forward image
    double2* d_vjp_cg = nullptr;   // VJP of a non-Parseval plan: [max_frames] (c_f, alpha_f)
    float* d_vjp_slab = nullptr;   // any-size VJP: [vjp_blocks][vjp_slab_floats] scratch
    int vjp_blocks = 0;
    int64_t vjp_slab_floats = 0;
    float* d_roww_nat = nullptr;   // specialised LGS plans: row operator in its natural [npx][L] layout
    const float* static_A = nullptr;   // registered pupil, statistics in d_amp_st / d_scale_st
    int64_t static_a_stride = 0;
    int static_frames = 0;
    double static_S0 = 0.0;
    double* d_amp_part = nullptr;  // [max_frames][amp_blocks][2]
    unsigned* d_counter = nullptr; // [max_frames]
    double* d_bpart = nullptr;     // [max_frames][n_parts]
    double* d_tpart = nullptr;     // [max_frames][n_parts][2] fused CoG frame-total partials
    shslopec_plan* cog = nullptr;  // bound ShSlopec plan of the fused forward (shwfs_plan_bind_slopec)
    int32_t* d_slot = nullptr;     // [n_sub] slope index per subaperture (-1: not valid)
    int32_t* d_order = nullptr;    // [n_sub] lit-first subaperture order of the static pupil (NGS MFMA, LGS)
    uint8_t* d_lit = nullptr;      // [n_sub] lit flags (scratch of the order computation)
    bool order_valid = false;      // d_order describes static_A
    float* d_cogw = nullptr;       // [2 npx] CoG x weights, y weights
    hipEvent_t ev_start = nullptr; // optional timing hook around the main kernel
    hipEvent_t ev_stop = nullptr;
};

struct shslopec_plan {
    int device = 0;
    int n_sub = 0, np_sub = 0, frame_h = 0, frame_w = 0, max_frames = 0;
    bool rows_path = false;        // cog_rows_kernel (separable weights, square blocks)
    bool vec4 = false;
    std::vector<int32_t> h_origin; // block origins (host copy; empty for index-list plans)
    std::vector<double> h_sep;     // xv | yv of separable weights of square blocks (empty otherwise)
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

// Optional timing hook of a plan (shwfs_plan_set_timing_events): the events are attached to the
// main kernel's dispatch (hipExtLaunchKernelGGL records them at the kernel's own start and end, so
// the interval is the kernel's duration, not that of separate stream markers around it).  Set by
// sh_forward_impl around the main launch only; null for every other launch.
thread_local hipEvent_t t_ev_start = nullptr, t_ev_stop = nullptr;

template <int NP, int M>
int launch_ngs(shwfs_plan* p, const NgsArgs& a, int n_frames, hipStream_t st) {
    constexpr int TS = 256 / NP;
    const int g1_bytes = TS * (M * NP + 1) * (int)sizeof(float2);
    const int u_bytes = TS * (p->npx_pad * M + 4) * (int)sizeof(float);
    const int lds = (g1_bytes > u_bytes ? g1_bytes : u_bytes);
    if (lds > 150 * 1024) return set_error(SHWFS_ERR_UNSUPPORTED, "LDS budget exceeded for this geometry");
    dim3 grid(p->n_tiles, n_frames);
    hipExtLaunchKernelGGL((sh_ngs_forward_kernel<NP, M, TS>), grid, dim3(TS * NP), lds, st, t_ev_start, t_ev_stop, 0, a);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

int ngs_tile_size(int np) { return 256 / np; }

template <int NP, bool PARSEVAL, int NPX, int COG, bool BF3>
int launch_mfma(const NgsArgs& a, int n_tiles, int n_frames, int fpb, hipStream_t st) {
    dim3 grid(n_tiles * NgsMfmaShape<NP>::WPB, (n_frames + fpb - 1) / fpb);
    hipExtLaunchKernelGGL((sh_ngs_mfma_kernel<NP, PARSEVAL, NPX, COG, BF3>), grid, dim3(64), 0, st, t_ev_start, t_ev_stop,
                          0, a, n_frames, fpb);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

// Fast CCD store layout for np 8 with 8 or 16 px per subaperture (aligned CCD), else the generic one.
inline int mfma_store_npx(int np, const NgsArgs& a) {
    if (np != 8 || !a.store_vec) return 0;
    return (a.npx == 16 || a.npx == 8) ? a.npx : 0;
}

template <int NP, bool BF3>
int launch_ngs_mfma_prec(const NgsArgs& a, int T, int n_frames, int fpb, hipStream_t st) {
    if constexpr (NP == 8) {
        const int npx = mfma_store_npx(8, a);
        if (a.cog) {   // fused CoG: Parseval scale and the fast store layout (checked at bind time)
            if (!a.parseval || npx == 0) return set_error(SHWFS_ERR_UNSUPPORTED, "fused CoG needs the fast Parseval path");
            const int mode = (a.cog == 1 && a.thr == 0.f) ? 3 : a.cog;   // 3: no threshold
            if (npx == 16) return mode == 3 ? launch_mfma<8, true, 16, 3, BF3>(a, T, n_frames, fpb, st)
                                : mode == 1 ? launch_mfma<8, true, 16, 1, BF3>(a, T, n_frames, fpb, st)
                                            : launch_mfma<8, true, 16, 2, BF3>(a, T, n_frames, fpb, st);
            return mode == 3 ? launch_mfma<8, true, 8, 3, BF3>(a, T, n_frames, fpb, st)
                 : mode == 1 ? launch_mfma<8, true, 8, 1, BF3>(a, T, n_frames, fpb, st)
                             : launch_mfma<8, true, 8, 2, BF3>(a, T, n_frames, fpb, st);
        }
        if (a.parseval) {
            if (npx == 16) return launch_mfma<8, true, 16, 0, BF3>(a, T, n_frames, fpb, st);
            if (npx == 8) return launch_mfma<8, true, 8, 0, BF3>(a, T, n_frames, fpb, st);
            return launch_mfma<8, true, 0, 0, BF3>(a, T, n_frames, fpb, st);
        }
        if (npx == 16) return launch_mfma<8, false, 16, 0, BF3>(a, T, n_frames, fpb, st);
        if (npx == 8) return launch_mfma<8, false, 8, 0, BF3>(a, T, n_frames, fpb, st);
        return launch_mfma<8, false, 0, 0, BF3>(a, T, n_frames, fpb, st);
    } else {
        if (a.cog) return set_error(SHWFS_ERR_UNSUPPORTED, "fused CoG needs np_sub 8");
        return a.parseval ? launch_mfma<NP, true, 0, 0, BF3>(a, T, n_frames, fpb, st)
                          : launch_mfma<NP, false, 0, 0, BF3>(a, T, n_frames, fpb, st);
    }
}

template <int NP>
int launch_ngs_mfma(shwfs_plan* p, const NgsArgs& a, int n_frames, hipStream_t st) {
    const int fpb = (n_frames == 1) ? 1 : p->fpb;
    return p->bf3 ? launch_ngs_mfma_prec<NP, true>(a, p->n_tiles, n_frames, fpb, st)
                  : launch_ngs_mfma_prec<NP, false>(a, p->n_tiles, n_frames, fpb, st);
}

using NgsLauncher = int (*)(shwfs_plan*, const NgsArgs&, int, hipStream_t);

template <int NP, int M, int TS>
int launch_lgs(shwfs_plan* p, const NgsArgs& a, int n_frames, hipStream_t st) {
    using Sh = LgsShape<NP, M, TS>;
    constexpr int lds = Sh::LDS_F2 * (int)sizeof(float2);
    static_assert(lds + Sh::COG_BYTES <= 64 * 1024, "LGS tile exceeds 64 KiB of LDS");
    dim3 grid(p->n_tiles, n_frames);
    const int mode = (a.cog == 1 && a.thr == 0.f) ? 3 : a.cog;   // ShSlopec moments in the epilogue
    if (mode == 3)
        hipExtLaunchKernelGGL((sh_lgs_forward_kernel<NP, M, TS, 3>), grid, dim3(Sh::BT), lds + Sh::COG_BYTES, st,
                              t_ev_start, t_ev_stop, 0, a);
    else if (mode == 1)
        hipExtLaunchKernelGGL((sh_lgs_forward_kernel<NP, M, TS, 1>), grid, dim3(Sh::BT), lds + Sh::COG_BYTES, st,
                              t_ev_start, t_ev_stop, 0, a);
    else if (mode == 2)
        hipExtLaunchKernelGGL((sh_lgs_forward_kernel<NP, M, TS, 2>), grid, dim3(Sh::BT), lds + Sh::COG_BYTES, st,
                              t_ev_start, t_ev_stop, 0, a);
    else
        hipExtLaunchKernelGGL((sh_lgs_forward_kernel<NP, M, TS, 0>), grid, dim3(Sh::BT), lds, st, t_ev_start,
                              t_ev_stop, 0, a);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

// [M/2] twiddles exp(-2 pi i k / M) of the any-size path's FFTs, evaluated in FP64
std::vector<float2> twiddle_table(int M) {
    std::vector<float2> t(M / 2);
    for (int k = 0; k < M / 2; ++k) {
        const double th = -2.0 * M_PI * (double)k / (double)M;
        t[k] = make_float2((float)std::cos(th), (float)std::sin(th));
    }
    return t;
}

int launch_general(shwfs_plan* p, const NgsArgs& a, int n_frames, hipStream_t st) {
    const shwfs_geom& g = p->g;
    GenArgs q{};
    q.tw = p->d_tw;
    q.slab = p->d_slab;
    q.slab_floats = p->slab_floats;
    q.M = g.spec_size;
    int lg = 0;
    while ((1 << lg) < q.M) ++lg;
    q.log2M = lg;
    q.np = g.np_sub;
    q.log2ch = gen_chunk_log2(q.M);             // vectors per chunk: rows padded to M + 1 in LDS
    q.ch = 1 << q.log2ch;
    q.terms = p->terms;
    q.term_stride = (int64_t)g.subap_npx * (g.lgs ? 2LL * p->L : (int64_t)g.spec_size);
    NgsArgs b = a;
    if (!g.lgs) b.kspec = nullptr;                 // the kernel takes the LGS branch on kspec != nullptr
    const int64_t items = (int64_t)n_frames * p->n_sub;
    const int blocks = (int)std::min<int64_t>(items, p->gen_blocks);
    if (gen_cap(q.M) == kGenLdsBigF2)
        hipExtLaunchKernelGGL(sh_general_kernel<kGenLdsBigF2>, dim3(blocks), dim3(kGenThreads), 0, st, t_ev_start,
                              t_ev_stop, 0, b, q, n_frames);
    else
        hipExtLaunchKernelGGL(sh_general_kernel<kGenLdsF2>, dim3(blocks), dim3(kGenThreads), 0, st, t_ev_start,
                              t_ev_stop, 0, b, q, n_frames);
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
    return shwfs_plan_create_masked(geom, 1, row_w, col_w, device, out);
}

int shwfs_plan_create_masked(const shwfs_geom* geom, int n_terms, const double* row_w, const double* col_w,
                             int device, shwfs_plan** out) {
    if (!geom || !row_w || !col_w || !out) return set_error(SHWFS_ERR_ARG, "null argument");
    const shwfs_geom& g = *geom;
    if (n_terms < 1 || n_terms > 4096) return set_error(SHWFS_ERR_ARG, "n_terms outside [1, 4096]");
    if (n_terms > 1 && g.variant != 0 && g.variant != 4)
        return set_error(SHWFS_ERR_ARG, "a non-separable FoV mask runs on the any-size path (variant 0 or 4)");
    if (n_terms > 1 && g.parseval)
        return set_error(SHWFS_ERR_ARG, "a non-separable FoV mask has no Parseval normalisation");
    if (g.max_frames > 65535) return set_error(SHWFS_ERR_ARG, "max_frames above 65535 (frames index the launch grid's y)");
    if (g.n_lenses <= 0 || g.np_sub <= 0 || g.subap_npx <= 0 || g.max_frames <= 0)
        return set_error(SHWFS_ERR_ARG, "non-positive geometry field");
    if (g.ef_size != g.n_lenses * g.np_sub)
        return set_error(SHWFS_ERR_ARG, "ef_size must equal n_lenses * np_sub");
    if (g.spec_size < 2 * g.np_sub - 1)
        return set_error(SHWFS_ERR_ARG, "spec_size must be >= 2*np_sub-1");
    if (g.variant < 0 || g.variant > 4) return set_error(SHWFS_ERR_ARG, "variant must be 0, 1, 2, 3 or 4");
    NgsLauncher mf = nullptr, v1 = nullptr, lg = nullptr;
    int lgs_ts = 0;
    bool use_mfma = false, use_general = g.variant == 4 || n_terms > 1;
    const bool general_fits = (g.spec_size & (g.spec_size - 1)) == 0 && g.spec_size >= 16 && g.spec_size <= 1024;
    if (g.lgs) {
        lg = pick_lgs(g.np_sub, g.spec_size, &lgs_ts);
        if (!lg || g.subap_npx > 16) use_general = true;     // no specialised LGS kernel: any-size path
        if (g.variant == 1 || g.variant == 2 || g.variant == 3)
            return set_error(SHWFS_ERR_ARG, "variants 1-3 are NGS binning variants");
    } else if (!use_general) {
        mf = pick_ngs_mfma(g.np_sub, g.spec_size, g.subap_npx);
        v1 = pick_ngs(g.np_sub, g.spec_size);
        use_mfma = (g.variant >= 2) || (g.variant == 0 && mf != nullptr);
        if (use_mfma && !mf) return set_error(SHWFS_ERR_UNSUPPORTED, "MFMA variant needs spec_size 16, np_sub 5..8, subap_npx <= 16");
        if (!use_mfma && !v1) {
            if (g.variant != 0) return set_error(SHWFS_ERR_UNSUPPORTED, "np_sub / spec_size combination not instantiated");
            use_general = true;                               // no specialised NGS kernel: any-size path
        }
    }
    if (use_general && !general_fits)
        return set_error(SHWFS_ERR_UNSUPPORTED, "spectrum grid must be a power of two in [16, 1024] (np_sub <= 512)");
    if (use_general && g.lgs && (int64_t)g.subap_npx * g.np_sub > kGenLdsF2 &&
        2LL * g.subap_npx * g.np_sub > (int64_t)g.spec_size * g.spec_size)
        return set_error(SHWFS_ERR_UNSUPPORTED, "LGS any-size path: subap_npx * np_sub too large for its work buffers");
    DeviceGuard dg(device);
    auto* p = new shwfs_plan();
    p->g = g;
    p->device = device;
    p->n_sub = g.n_lenses * g.n_lenses;
    p->npx_pad = (g.subap_npx + 3) / 4 * 4;
    p->mfma = use_mfma;
    p->general = use_general;
    p->terms = n_terms;
    p->bf3 = use_mfma && g.variant == 3;
    p->L = 2 * g.np_sub - 1;
    p->amp_blocks = 256;
    const int M = g.spec_size;
    std::vector<float> rw, cw;
    if (use_general) {
        // operators in their natural row-major layout: NGS [npx][M] real, LGS [npx][L] complex
        const size_t per = g.lgs ? (size_t)p->L * 2 : (size_t)M;
        p->rows_alloc = g.subap_npx;
        rw.assign((size_t)n_terms * g.subap_npx * per, 0.f);
        cw.assign((size_t)n_terms * g.subap_npx * per, 0.f);
        for (size_t q = 0; q < rw.size(); ++q) { rw[q] = (float)row_w[q]; cw[q] = (float)col_w[q]; }
        p->ts = 1;
    } else if (g.lgs) {
        p->rows_alloc = 16;
        p->ts = lgs_ts;
        rw.assign((size_t)16 * p->L * 2, 0.f);
        cw.assign((size_t)16 * p->L * 2, 0.f);
        for (size_t q = 0; q < (size_t)g.subap_npx * p->L * 2; ++q) cw[q] = (float)col_w[q];
        // P transposed, [my][Y (16)][re, im], times M: the kernel's lag rows are the x
        // autocorrelations without the factor M of FFT_M(|FFT_M(row)|^2)
        for (int Y = 0; Y < g.subap_npx; ++Y)
            for (int mi = 0; mi < p->L; ++mi)
                for (int c = 0; c < 2; ++c)
                    rw[((size_t)mi * 16 + Y) * 2 + c] = (float)(row_w[((size_t)Y * p->L + mi) * 2 + c] * M);
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
    p->n_parts = (!g.lgs && use_mfma) ? p->n_tiles * NgsMfmaShape<8>::WPB : p->n_tiles;
    if (use_general) {
        // per-block scratch (sh_general.h): G / T, then U (NGS) or S and D (LGS); the number of
        // persistent blocks keeps the total at or below 1 GiB
        const int64_t gt = 2LL * g.np_sub * M;
        const int64_t bacc = n_terms > 1 ? (int64_t)g.subap_npx * g.subap_npx : 0;   // K-term accumulator
        p->slab_floats = g.lgs ? gt + (int64_t)M * M + 2LL * p->L * g.np_sub + bacc
                               : gt + (n_terms > 1 ? (int64_t)M * M : 0) + (int64_t)g.subap_npx * M + bacc;
        p->slab_floats = (p->slab_floats + 63) / 64 * 64;          // 256-byte aligned slabs
        const int64_t cap = ((int64_t)1 << 30) / (p->slab_floats * (int64_t)sizeof(float));
        p->gen_blocks = (int)std::max<int64_t>(1, std::min<int64_t>(1024, cap));
    }
    auto cleanup = [&](int code) { shwfs_plan_destroy(p); return code; };
    hipError_t e = hipSuccess;
#define PLAN_ALLOC(ptr, bytes)                                                                  \
    e = hipMalloc((void**)&(ptr), (bytes));                                                     \
    if (e != hipSuccess) return cleanup(set_error(SHWFS_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e)));
    PLAN_ALLOC(p->d_roww, rw.size() * sizeof(float));
    std::vector<float> rl;
    if (!g.lgs && use_mfma) {
        // P~[Y][j]: j = 0 -> C[0]; j = d -> Re C[d]; j = np-1+d -> Im C[d]   (d = 1 .. np-1), with
        // |G[ky]|^2 = C[0] + 2 sum_d (Re C[d] cos(2 pi ky d/M) + Im C[d] sin(2 pi ky d/M))
        rl.assign((size_t)16 * M, 0.f);
        for (int Y = 0; Y < g.subap_npx; ++Y) {
            const double* pr = row_w + (size_t)Y * M;
            double c0 = 0.0;
            for (int ky = 0; ky < M; ++ky) c0 += pr[ky];
            rl[(size_t)Y * M] = (float)c0;
            for (int d = 1; d < g.np_sub; ++d) {
                double cr = 0.0, ci = 0.0;
                for (int ky = 0; ky < M; ++ky) {
                    const double th = 2.0 * M_PI * (double)((ky * d) % M) / M;
                    cr += 2.0 * pr[ky] * std::cos(th);
                    ci += 2.0 * pr[ky] * std::sin(th);
                }
                rl[(size_t)Y * M + d] = (float)cr;
                rl[(size_t)Y * M + g.np_sub - 1 + d] = (float)ci;
            }
        }
        PLAN_ALLOC(p->d_rowlag, rl.size() * sizeof(float));
    }
    PLAN_ALLOC(p->d_colw, cw.size() * sizeof(float));
    PLAN_ALLOC(p->d_amp, (size_t)g.max_frames * 2 * sizeof(double));
    PLAN_ALLOC(p->d_scale, (size_t)g.max_frames * sizeof(float));
    PLAN_ALLOC(p->d_amp_st, (size_t)g.max_frames * 2 * sizeof(double));
    PLAN_ALLOC(p->d_scale_st, (size_t)g.max_frames * sizeof(float));
    PLAN_ALLOC(p->d_amp_part, (size_t)g.max_frames * p->amp_blocks * 2 * sizeof(double));
    PLAN_ALLOC(p->d_counter, (size_t)g.max_frames * sizeof(unsigned));
    PLAN_ALLOC(p->d_bpart, (size_t)g.max_frames * p->n_parts * sizeof(double));
    if (g.lgs || use_mfma) {
        PLAN_ALLOC(p->d_order, (size_t)p->n_sub * sizeof(int32_t));
        PLAN_ALLOC(p->d_lit, (size_t)p->n_sub);
    }
    if (g.lgs) PLAN_ALLOC(p->d_kspec, (size_t)p->n_sub * p->L * p->L * 2 * sizeof(float));
    std::vector<float> rnat;
    if (g.lgs && !use_general) {        // the any-size VJP reads P~ in its natural layout
        rnat.resize((size_t)g.subap_npx * p->L * 2);
        for (size_t q = 0; q < rnat.size(); ++q) rnat[q] = (float)row_w[q];
        PLAN_ALLOC(p->d_roww_nat, rnat.size() * sizeof(float));
    }
    std::vector<float2> twh;
    if (use_general) {
        PLAN_ALLOC(p->d_tw, (size_t)(M / 2) * sizeof(float2));
        PLAN_ALLOC(p->d_slab, (size_t)p->gen_blocks * p->slab_floats * sizeof(float));
        twh = twiddle_table(M);
    }
#undef PLAN_ALLOC
    if ((!rl.empty() &&
         (e = hipMemcpy(p->d_rowlag, rl.data(), rl.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) ||
        (e = hipMemcpy(p->d_roww, rw.data(), rw.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
        (!rnat.empty() &&
         (e = hipMemcpy(p->d_roww_nat, rnat.data(), rnat.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) ||
        (e = hipMemcpy(p->d_colw, cw.data(), cw.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
        (!twh.empty() &&
         (e = hipMemcpy(p->d_tw, twh.data(), twh.size() * sizeof(float2), hipMemcpyHostToDevice)) != hipSuccess) ||
        (e = hipMemset(p->d_counter, 0, (size_t)g.max_frames * sizeof(unsigned))) != hipSuccess)
        return cleanup(set_error(SHWFS_ERR_HIP, std::string("plan upload: ") + hipGetErrorString(e)));
    *out = p;
    return SHWFS_OK;
}

int shwfs_plan_destroy(shwfs_plan* p) {
    if (!p) return SHWFS_OK;
    DeviceGuard dg(p->device);
    (void)hipFree(p->d_roww);
    (void)hipFree(p->d_rowlag);
    (void)hipFree(p->d_colw);
    (void)hipFree(p->d_kspec);
    (void)hipFree(p->d_amp);
    (void)hipFree(p->d_scale);
    (void)hipFree(p->d_amp_st);
    (void)hipFree(p->d_scale_st);
    (void)hipFree(p->d_amp_part);
    (void)hipFree(p->d_counter);
    (void)hipFree(p->d_bpart);
    (void)hipFree(p->d_slot);
    (void)hipFree(p->d_order);
    (void)hipFree(p->d_lit);
    (void)hipFree(p->d_tpart);
    (void)hipFree(p->d_ext);
    (void)hipFree(p->d_A_rs);
    (void)hipFree(p->d_A_rs_st);
    (void)hipFree(p->d_ph_rs);
    (void)hipFree(p->d_gph_rs);
    (void)hipFree(p->d_vjp_ccd);
    (void)hipFree(p->d_vjp_cg);
    (void)hipFree(p->d_vjp_slab);
    (void)hipFree(p->d_roww_nat);
    (void)hipFree(p->d_cogw);
    (void)hipFree(p->d_tw);
    (void)hipFree(p->d_slab);
    delete p;
    return SHWFS_OK;
}

int shwfs_plan_set_frames_per_block(shwfs_plan* p, int fpb) {
    if (!p) return set_error(SHWFS_ERR_ARG, "null plan");
    if (fpb < 1 || fpb > 64) return set_error(SHWFS_ERR_ARG, "frames per block must be in [1, 64]");
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

}  // extern "C"

namespace {
inline bool aligned16(const void* ptr) { return (reinterpret_cast<uintptr_t>(ptr) & 15u) == 0; }

// Fused CoG request of one forward call (nullptr: plain SH forward).
struct CogCall {
    float thr;
    int thr_mode;
};

int launch_resample(const shwfs_plan* p, const float* in, int64_t in_stride, const int2* ext, float* out,
                    int n_frames, hipStream_t st) {
    const int64_t n_out = (int64_t)p->g.ef_size * p->g.ef_size;
    const int blocks = (int)std::min<int64_t>((n_out + 255) / 256, 2048);
    hipLaunchKernelGGL(resample_kernel, dim3(blocks, n_frames), dim3(256), 0, st, in, in_stride, ext, out, n_out,
                       p->ef_in, p->g.ef_size, p->rmap);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

// Statistics of the amplitude frames (masked area, square sum, Parseval scale) into amp / scale;
// with resampling the amplitude is first resampled (d_A_rs) and the square sum is taken there.
int pupil_stats(shwfs_plan* p, const float* A, int64_t a_stride, int amp_frames, double S0, double* amp,
                float* scale, float* A_rs, hipStream_t st) {
    const shwfs_geom& g = p->g;
    const int64_t ef_px = (int64_t)g.ef_size * g.ef_size;
    const float* B = nullptr;
    int64_t b_stride = 0, n_in = ef_px;
    if (p->ef_in) {
        n_in = (int64_t)p->ef_in * p->ef_in;
        int rc = launch_resample(p, A, a_stride, nullptr, A_rs, amp_frames, st);
        if (rc) return rc;
        B = A_rs;
        b_stride = ef_px;
    }
    hipLaunchKernelGGL(amp_stats_kernel, dim3(p->amp_blocks, amp_frames), dim3(256), 0, st, A, a_stride, n_in,
                       p->d_amp_part, p->d_counter, amp, S0 * g.pixel_area, g.norm_const,
                       g.parseval ? scale : nullptr, B, b_stride, ef_px);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

int sh_forward_impl(shwfs_plan* p, int n_frames, const float* A, int64_t a_stride, const float* phase,
                    int64_t ph_stride, double S0, float* ccd, int64_t ccd_stride, const CogCall* cog,
                    hipStream_t st) {
    if (!p || !A || !phase || !ccd) return set_error(SHWFS_ERR_ARG, "null argument");
    if (n_frames <= 0 || n_frames > p->g.max_frames)
        return set_error(SHWFS_ERR_ARG, "n_frames outside [1, max_frames]");
    const shwfs_geom& g = p->g;
    const int64_t ef_px = (int64_t)g.ef_size * g.ef_size;
    const int64_t in_px = p->ef_in ? (int64_t)p->ef_in * p->ef_in : ef_px;   // input field size
    const int64_t ccd_side = (int64_t)g.n_lenses * g.subap_npx;
    if (a_stride != 0 && a_stride < in_px) return set_error(SHWFS_ERR_ARG, "A frame stride too small");
    if (n_frames > 1 && ph_stride != 0 && ph_stride < in_px) return set_error(SHWFS_ERR_ARG, "phase frame stride too small");
    if (n_frames > 1 && ccd_stride < ccd_side * ccd_side)
        return set_error(SHWFS_ERR_ARG, "ccd frame stride too small");
    DeviceGuard dg(p->device);

    // amplitude statistics: one pass per distinct amplitude frame, unless this is the pupil
    // registered with shwfs_plan_set_pupil (statistics already on the device, in buffers of their
    // own, so that calls with another amplitude or S0 leave the registration intact)
    const int amp_frames = (a_stride == 0) ? 1 : n_frames;
    const bool cached = A == p->static_A && a_stride == p->static_a_stride && S0 == p->static_S0 &&
                        amp_frames <= p->static_frames;
    double* amp = cached ? p->d_amp_st : p->d_amp;
    float* fscale = cached ? p->d_scale_st : p->d_scale;
    if (!cached) {
        int rc = pupil_stats(p, A, a_stride, amp_frames, S0, amp, fscale, p->d_A_rs, st);
        if (rc) return rc;
    }
    if (p->ef_in) {   // resampled field: the SH kernels read the resampled amplitude and d_ph_rs
        int rc = launch_resample(p, phase, ph_stride, p->d_ext, p->d_ph_rs, n_frames, st);
        if (rc) return rc;
        A = cached ? p->d_A_rs_st : p->d_A_rs;
        a_stride = (a_stride == 0) ? 0 : ef_px;
        phase = p->d_ph_rs;
        ph_stride = ef_px;
    }

    NgsArgs a{};
    a.fscale = fscale; a.scale_fstep = (a_stride == 0) ? 0 : 1;
    a.A = A; a.a_stride = a_stride;
    a.phase = phase; a.ph_stride = ph_stride;
    a.ccd = ccd; a.ccd_stride = ccd_stride;
    a.roww = (p->mfma && !g.lgs) ? p->d_rowlag : p->d_roww; a.colw = p->d_colw;
    a.amp = amp; a.amp_fstep = (a_stride == 0) ? 0 : 2;
    a.bpart = p->d_bpart; a.n_parts = cog ? p->n_cog_parts : p->n_parts;
    a.n_lenses = g.n_lenses; a.ef_size = g.ef_size; a.npx = g.subap_npx; a.npx_pad = p->npx_pad;
    a.n_sub = p->n_sub; a.n_tiles = p->n_tiles; a.parseval = g.parseval;
    a.k = g.k_rad_per_nm; a.tilt = g.tilt_rad; a.a0 = g.tilt_a0;
    a.S0 = S0; a.pixel_area = g.pixel_area; a.norm_const = g.norm_const;
    a.kspec = reinterpret_cast<const float2*>(p->d_kspec);
    a.vec = (g.ef_size % 4 == 0) && aligned16(A) && aligned16(phase) && (a_stride % 4 == 0) &&
            (n_frames == 1 || ph_stride % 4 == 0);
    a.store_vec = aligned16(ccd) && (ccd_side % 4 == 0) && (n_frames == 1 || ccd_stride % 4 == 0);
    // lit-first order of the registered pupil (Parseval only: the images do not depend on how
    // subapertures are grouped into waves, the non-Parseval partial sums would)
    a.order = (cached && p->order_valid && g.parseval) ? p->d_order : nullptr;
    if (cog) {
        a.cog = cog->thr_mode + 1;
        a.thr = cog->thr;
        a.slot = p->d_slot;
        a.mom = p->cog->d_mom;
        a.cogw = p->d_cogw;
        a.tpart = p->d_tpart;
        a.n_valid = p->cog->n_sub;
    }
    if (g.lgs && !p->kernels_set) return set_error(SHWFS_ERR_ARG, "LGS plan: call shwfs_set_kernels first");
    int lts = 0;
    NgsLauncher L = p->general ? launch_general
                  : g.lgs ? pick_lgs(g.np_sub, g.spec_size, &lts)
                          : (p->mfma ? pick_ngs_mfma(g.np_sub, g.spec_size, g.subap_npx) : pick_ngs(g.np_sub, g.spec_size));
    t_ev_start = p->ev_start;
    t_ev_stop = p->ev_stop;
    int rc = L(p, a, n_frames, st);
    t_ev_start = t_ev_stop = nullptr;
    if (rc) return rc;
    if (!g.parseval) {
        hipLaunchKernelGGL(sh_rescale_kernel, dim3(64, n_frames), dim3(256), 0, st, ccd, ccd_stride,
                           ccd_side * ccd_side, p->d_bpart, p->n_parts, amp,
                           (a_stride == 0) ? 0 : 2, S0, g.pixel_area);
        SHWFS_LAUNCH_CHECK();
    }
    return SHWFS_OK;
}
}  // namespace

extern "C" {

int shwfs_forward(shwfs_plan* p, int n_frames, const float* A, int64_t a_stride,
                  const float* phase, int64_t ph_stride, double S0, float* ccd,
                  int64_t ccd_stride, void* stream) {
    return sh_forward_impl(p, n_frames, A, a_stride, phase, ph_stride, S0, ccd, ccd_stride, nullptr,
                           reinterpret_cast<hipStream_t>(stream));
}

int shwfs_plan_set_pupil(shwfs_plan* p, const float* A, int64_t a_stride, int n_frames, double S0, void* stream) {
    if (!p) return set_error(SHWFS_ERR_ARG, "null plan");
    p->static_A = nullptr;
    if (!A) return SHWFS_OK;
    const shwfs_geom& g = p->g;
    const int64_t in_px = p->ef_in ? (int64_t)p->ef_in * p->ef_in : (int64_t)g.ef_size * g.ef_size;
    if (n_frames <= 0 || n_frames > g.max_frames) return set_error(SHWFS_ERR_ARG, "n_frames outside [1, max_frames]");
    if (a_stride != 0 && a_stride < in_px) return set_error(SHWFS_ERR_ARG, "A frame stride too small");
    const int amp_frames = (a_stride == 0) ? 1 : n_frames;
    DeviceGuard dg(p->device);
    if (p->ef_in && p->A_rs_st_frames < amp_frames) {   // set-up time allocation of the resampled copy
        (void)hipFree(p->d_A_rs_st);
        p->d_A_rs_st = nullptr;
        p->A_rs_st_frames = 0;
        SHWFS_HIP_CHECK(hipMalloc((void**)&p->d_A_rs_st, (size_t)amp_frames * g.ef_size * g.ef_size * sizeof(float)));
        p->A_rs_st_frames = amp_frames;
    }
    int rc = pupil_stats(p, A, a_stride, amp_frames, S0, p->d_amp_st, p->d_scale_st, p->d_A_rs_st,
                         reinterpret_cast<hipStream_t>(stream));
    if (rc) return rc;
    p->order_valid = false;
    if (p->d_order && a_stride == 0 && !p->ef_in) {
        rc = launch_subap_order(A, g.ef_size, g.n_lenses, g.np_sub, p->n_sub, p->d_lit, p->d_order,
                                reinterpret_cast<hipStream_t>(stream));
        if (rc) return rc;
        p->order_valid = true;
    }
    p->static_A = A;
    p->static_a_stride = a_stride;
    p->static_frames = amp_frames;
    p->static_S0 = S0;
    return SHWFS_OK;
}

int shwfs_plan_pupil_registered(const shwfs_plan* p, const float* A, int64_t a_stride, double S0) {
    if (!p) return set_error(SHWFS_ERR_ARG, "null plan");
    return (A != nullptr && A == p->static_A && a_stride == p->static_a_stride && S0 == p->static_S0) ? 1 : 0;
}

int shwfs_plan_set_resampling(shwfs_plan* p, int ef_in, const int32_t* ext_src1, const int32_t* ext_src2,
                              const double* affine) {
    if (!p || !affine) return set_error(SHWFS_ERR_ARG, "null argument");
    if (ef_in < 2) return set_error(SHWFS_ERR_ARG, "input field side must be >= 2");
    if ((ext_src1 == nullptr) != (ext_src2 == nullptr)) return set_error(SHWFS_ERR_ARG, "give both source maps or neither");
    const int64_t n_in = (int64_t)ef_in * ef_in;
    std::vector<int2> ext;
    if (ext_src1) {
        ext.resize(n_in);
        for (int64_t q = 0; q < n_in; ++q) {   // host-side bounds check of every gather index
            if (ext_src1[q] < -1 || ext_src1[q] >= n_in || ext_src2[q] < -1 || ext_src2[q] >= n_in ||
                (ext_src1[q] < 0 && ext_src2[q] >= 0))
                return set_error(SHWFS_ERR_ARG, "edge-extrapolation source index out of range");
            ext[q] = make_int2(ext_src1[q], ext_src2[q]);
        }
    }
    const shwfs_geom& g = p->g;
    const size_t rs_bytes = (size_t)g.max_frames * g.ef_size * g.ef_size * sizeof(float);
    DeviceGuard dg(p->device);
    (void)hipFree(p->d_ext); p->d_ext = nullptr;
    if (!p->d_A_rs) SHWFS_HIP_CHECK(hipMalloc((void**)&p->d_A_rs, rs_bytes));
    if (!p->d_ph_rs) SHWFS_HIP_CHECK(hipMalloc((void**)&p->d_ph_rs, rs_bytes));
    if (!ext.empty()) {
        SHWFS_HIP_CHECK(hipMalloc((void**)&p->d_ext, n_in * sizeof(int2)));
        SHWFS_HIP_CHECK(hipMemcpy(p->d_ext, ext.data(), n_in * sizeof(int2), hipMemcpyHostToDevice));
    }
    p->rmap = ResampleMap{affine[0], affine[1], affine[2], affine[3], affine[4], affine[5]};
    p->ef_in = ef_in;
    p->static_A = nullptr;
    return SHWFS_OK;
}

int shwfs_plan_bind_slopec(shwfs_plan* p, shslopec_plan* c) {
    if (!p) return set_error(SHWFS_ERR_ARG, "null plan");
    DeviceGuard dg(p->device);
    p->cog = nullptr;
    if (!c) return SHWFS_OK;
    const shwfs_geom& g = p->g;
    const int n = g.n_lenses, npx = g.subap_npx, W = n * npx;
    int lgs_ts = 0;
    const bool lgs_ok = g.lgs && !p->general && pick_lgs(g.np_sub, g.spec_size, &lgs_ts) && npx <= 16;
    const bool ngs_ok = !g.lgs && p->mfma && g.np_sub == 8 && (npx == 8 || npx == 16);
    if (!g.parseval || !(lgs_ok || ngs_ok))
        return set_error(SHWFS_ERR_UNSUPPORTED, "fused CoG needs Parseval normalisation and an NGS MFMA plan with "
                                                "np_sub 8 and 8 or 16 px per subaperture, or an LGS plan on a "
                                                "specialised LGS kernel");
    if (c->device != p->device) return set_error(SHWFS_ERR_ARG, "plans on different devices");
    if (c->h_sep.empty() || c->np_sub != npx || c->frame_h != W || c->frame_w != W)
        return set_error(SHWFS_ERR_UNSUPPORTED, "fused CoG needs square subaperture blocks of the SH pixel "
                                                "grid with separable weights (the ShSlopec default)");
    std::vector<int32_t> slot((size_t)p->n_sub, -1);
    for (int s = 0; s < c->n_sub; ++s) {
        const int32_t o = c->h_origin[s];
        const int r0 = o / W, c0 = o % W;
        if (r0 % npx != 0 || c0 % npx != 0)
            return set_error(SHWFS_ERR_UNSUPPORTED, "subaperture block not aligned to the lenslet grid");
        const int sub = (r0 / npx) * n + c0 / npx;
        if (slot[sub] >= 0) return set_error(SHWFS_ERR_UNSUPPORTED, "two subapertures share one lenslet block");
        slot[sub] = s;
    }
    std::vector<float> w(2 * (size_t)npx);
    for (int k = 0; k < 2 * npx; ++k) w[k] = (float)c->h_sep[k];
    if (!p->d_slot) SHWFS_HIP_CHECK(hipMalloc((void**)&p->d_slot, (size_t)p->n_sub * sizeof(int32_t)));
    if (!p->d_cogw) SHWFS_HIP_CHECK(hipMalloc((void**)&p->d_cogw, 2 * (size_t)npx * sizeof(float)));
    p->n_cog_parts = g.lgs ? p->n_tiles * lgs_ts : p->n_parts;
    if (!p->d_tpart)
        SHWFS_HIP_CHECK(hipMalloc((void**)&p->d_tpart, (size_t)g.max_frames * p->n_cog_parts * 2 * sizeof(double)));
    SHWFS_HIP_CHECK(hipMemcpy(p->d_slot, slot.data(), slot.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    SHWFS_HIP_CHECK(hipMemcpy(p->d_cogw, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
    p->cog = c;
    return SHWFS_OK;
}

int shwfs_forward_slopes(shwfs_plan* p, int n_frames, const float* A, int64_t a_stride,
                         const float* phase, int64_t ph_stride, double S0, float* ccd, int64_t ccd_stride,
                         float thr, int thr_mode, float* slopes, int64_t slopes_stride, double* flux,
                         double* totals, void* stream) {
    if (!p) return set_error(SHWFS_ERR_ARG, "null plan");
    shslopec_plan* c = p->cog;
    if (!c) return set_error(SHWFS_ERR_ARG, "no ShSlopec plan bound (shwfs_plan_bind_slopec)");
    if (!slopes || !flux || !totals) return set_error(SHWFS_ERR_ARG, "null argument");
    if (thr_mode != 0 && thr_mode != 1) return set_error(SHWFS_ERR_ARG, "thr_mode must be 0 or 1");
    if (n_frames <= 0 || n_frames > c->max_frames)
        return set_error(SHWFS_ERR_ARG, "n_frames outside [1, max_frames] of the ShSlopec plan");
    if (n_frames > 1 && slopes_stride < 2 * (int64_t)c->n_sub)
        return set_error(SHWFS_ERR_ARG, "slopes frame stride too small");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int64_t ccd_side = (int64_t)p->g.n_lenses * p->g.subap_npx;
    const bool fast = aligned16(ccd) && (n_frames == 1 || ccd_stride % 4 == 0) && (ccd_side % 4 == 0);
    if (!fast) {   // unaligned CCD buffer: the two passes give the same result
        int rc = sh_forward_impl(p, n_frames, A, a_stride, phase, ph_stride, S0, ccd, ccd_stride, nullptr, st);
        if (rc) return rc;
        return shslopec_forward(c, n_frames, ccd, ccd_stride, thr, thr_mode, slopes, slopes_stride, flux,
                                totals, stream);
    }
    const CogCall cc{thr, thr_mode};
    int rc = sh_forward_impl(p, n_frames, A, a_stride, phase, ph_stride, S0, ccd, ccd_stride, &cc, st);
    if (rc) return rc;
    DeviceGuard dg(p->device);
    hipLaunchKernelGGL(cog_finalize_parts_kernel, dim3((c->n_sub + 1023) / 1024, n_frames), dim3(256), 0, st,
                       c->d_mom, p->d_tpart, p->n_cog_parts, c->n_sub, slopes, slopes_stride, flux, totals,
                       (double*)nullptr);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

int shslopec_plan_create(int n_sub, int np_sub, int frame_h, int frame_w,
                         const int32_t* block_origin, const int32_t* idx,
                         const double* xw, const double* yw, const double* mw,
                         int max_frames, int device, shslopec_plan** out) {
    if (!out || !xw || !yw || !mw) return set_error(SHWFS_ERR_ARG, "null argument");
    if (max_frames > 65535) return set_error(SHWFS_ERR_ARG, "max_frames above 65535 (frames index the launch grid's y)");
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
    if (block_origin) p->h_origin.assign(block_origin, block_origin + n_sub);
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
        // separable default weights (x weight by column, y weight by row, no mask): kept for the
        // fused CoG of shwfs_plan_bind_slopec; with 4, 8 or 16 px blocks also the row kernel with
        // row-major processing order
        bool sep = true;
        for (int64_t r = 0; sep && r < np_sub; ++r)
            for (int64_t c = 0; c < np_sub; ++c) {
                const int64_t q = r * np_sub + c;
                if (mw[q] != 1.0 || xw[q] != xw[c] || yw[q] != yw[r * np_sub]) { sep = false; break; }
            }
        if (sep) {
            p->h_sep.resize(2 * (size_t)np_sub);
            for (int c = 0; c < np_sub; ++c) p->h_sep[c] = xw[c];
            for (int r = 0; r < np_sub; ++r) p->h_sep[np_sub + r] = yw[(int64_t)r * np_sub];
        }
        if (sep && (np_sub == 4 || np_sub == 8 || np_sub == 16)) {
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
template <int NP, int M, int TS, int NPX>
int launch_vjp(const VjpArgs& a, int n_tiles, int n_frames, int npx, hipStream_t st) {
    if (NPX > 0 && npx != NPX) return set_error(SHWFS_ERR_ARG, "VJP instantiation does not match subap_npx");
    const int lds = TS * (M * NP + 1) * (int)sizeof(float2) + (TS * npx * npx + 2 * npx * M) * (int)sizeof(float);
    if (lds > 150 * 1024) return set_error(SHWFS_ERR_UNSUPPORTED, "VJP tile exceeds the LDS budget");
    hipLaunchKernelGGL((sh_ngs_vjp_kernel<NP, M, TS, NPX>), dim3(n_tiles, n_frames), dim3(TS * NP), lds, st, a);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}
using VjpLauncher = int (*)(const VjpArgs&, int, int, int, hipStream_t);
VjpLauncher pick_vjp(int np, int M, int npx, int* ts) {
    if (np == 8 && M == 16) {
        *ts = 16;
        if (npx == 16) return launch_vjp<8, 16, 16, 16>;
        if (npx == 8) return launch_vjp<8, 16, 16, 8>;
        return launch_vjp<8, 16, 16, 0>;
    }
    if (np == 6 && M == 16) { *ts = 42; return launch_vjp<6, 16, 42, 0>; }
    if (np == 4 && M == 8) { *ts = 64; return launch_vjp<4, 8, 64, 0>; }
    if (np == 16 && M == 32) { *ts = 16; return launch_vjp<16, 32, 16, 0>; }
    return nullptr;
}
// ---------------------------------------------------------------- interaction-matrix launches
// Argument checks shared by shwfs_im_from_jacobian and its no-allocation variant.
constexpr int kImModesPerBlock = 16;

int im_check_args(int n_lenses, int np_sub, int ef_size, const float* jx, const float* jy, int n_modes,
                  const float* modes, int64_t mode_stride, const int32_t* slot, int n_valid, float* im) {
    if (!jx || !jy || !modes || !slot || !im) return set_error(SHWFS_ERR_ARG, "null argument");
    if (n_lenses <= 0 || np_sub <= 0 || n_modes <= 0 || n_valid <= 0 || ef_size != n_lenses * np_sub)
        return set_error(SHWFS_ERR_ARG, "inconsistent sizes");
    if (mode_stride < (int64_t)ef_size * ef_size) return set_error(SHWFS_ERR_ARG, "mode stride too small");
    if (n_modes > 65535) return set_error(SHWFS_ERR_ARG, "more than 65535 modes in one call");
    return SHWFS_OK;
}

// np_sub 4, 8, 16 on pupils up to 1024 px wide: the row-strip kernels; anything else the any-size one
bool im_strip_kernels(int np_sub, int ef_size) {
    return (np_sub == 4 || np_sub == 8 || np_sub == 16) && ef_size <= 1024;
}

int im_launch_any(int n_lenses, int np_sub, int ef_size, const float* jx, const float* jy, int n_modes,
                  const float* modes, int64_t mode_stride, const int32_t* slot, int n_valid, float scale, float* im,
                  hipStream_t st) {
    dim3 grid(n_lenses * n_lenses, (n_modes + kImModesPerBlock - 1) / kImModesPerBlock);
    hipLaunchKernelGGL(im_from_jacobian_any_kernel, grid, dim3(256), 0, st, jx, jy, ef_size, n_lenses, np_sub, modes,
                       mode_stride, n_modes, kImModesPerBlock, slot, n_valid, scale, im);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

bool im_v4_eligible(int np_sub, int ef_size, int64_t mode_stride, const float* modes, const float* jx,
                    const float* jy) {
    return (np_sub == 8 || np_sub == 16) && ef_size % 4 == 0 && mode_stride % 4 == 0 &&
           ((reinterpret_cast<uintptr_t>(modes) | reinterpret_cast<uintptr_t>(jx) |
             reinterpret_cast<uintptr_t>(jy)) & 15u) == 0;
}

// Scalar contraction, results written in slot order (no scratch).
int im_launch_scalar(int n_lenses, int np_sub, int ef_size, const float* jx, const float* jy, int n_modes,
                     const float* modes, int64_t mode_stride, const int32_t* slot, int n_valid, float scale, float* im,
                     hipStream_t st) {
    dim3 grid(n_lenses, (n_modes + kImModesPerBlock - 1) / kImModesPerBlock);
    const int threads = (ef_size + 63) / 64 * 64;
#define IMJ(NPV) hipLaunchKernelGGL(im_from_jacobian_kernel<NPV>, grid, dim3(threads), 0, st, jx, jy, ef_size, \
                                    n_lenses, modes, mode_stride, n_modes, kImModesPerBlock, slot, n_valid, scale, im)
    switch (np_sub) {
        case 4: IMJ(4); break;
        case 8: IMJ(8); break;
        default: IMJ(16); break;
    }
#undef IMJ
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

// float4 contraction.  gridbuf == nullptr: results written in slot order directly.  Otherwise the
// modes go in chunks of `chunk`: each chunk's results are written in grid order into gridbuf
// (chunk * 2 * n_lenses^2 floats, coalesced) and permuted into slot order by a gather pass through
// inv (n_valid int32, built once).
int im_launch_v4(int n_lenses, int np_sub, int ef_size, const float* jx, const float* jy, int n_modes,
                 const float* modes, int64_t mode_stride, const int32_t* slot, int n_valid, float scale, float* im,
                 float* gridbuf, int32_t* inv, int chunk, hipStream_t st) {
    const int threads4 = (ef_size / 4 * (np_sub / 4) + 63) / 64 * 64;
    const int nn = n_lenses * n_lenses;
    const bool grid_order = gridbuf != nullptr;
    if (grid_order) {
        hipLaunchKernelGGL(slot_inverse_kernel, dim3(1), dim3(1024), 0, st, slot, nn, n_valid, inv);
        SHWFS_LAUNCH_CHECK();
    } else {
        chunk = n_modes;
    }
    for (int m0 = 0; m0 < n_modes; m0 += chunk) {
        const int c = std::min(chunk, n_modes - m0);
        const float* md = modes + (int64_t)m0 * mode_stride;
        float* im_c = im + (int64_t)m0 * 2 * n_valid;
        dim3 grid(n_lenses, (c + kImModesPerBlock - 1) / kImModesPerBlock);
#define IMJ4(NPV, G) hipLaunchKernelGGL((im_from_jacobian_v4_kernel<NPV, G>), grid, dim3(threads4), 0, st, jx, jy,     \
                                        ef_size, n_lenses, md, mode_stride, c, kImModesPerBlock, slot, n_valid, scale, \
                                        G ? gridbuf : im_c)
        if (np_sub == 8) {
            if (grid_order) IMJ4(8, true); else IMJ4(8, false);
        } else {
            if (grid_order) IMJ4(16, true); else IMJ4(16, false);
        }
#undef IMJ4
        SHWFS_LAUNCH_CHECK();
        if (grid_order) {
            const int lds = 2 * nn * (int)sizeof(float);
            if (lds <= 64 * 1024)
                hipLaunchKernelGGL(im_gather_lds_kernel, dim3(c), dim3(256), lds, st, gridbuf, inv, nn, n_valid, im_c);
            else
                hipLaunchKernelGGL(im_gather_kernel, dim3((n_valid + 255) / 256, c), dim3(256), 0, st, gridbuf, inv,
                                   nn, n_valid, im_c);
            SHWFS_LAUNCH_CHECK();
        }
    }
    return SHWFS_OK;
}

// ---------------------------------------------------------------- no-allocation scratch
// Scratch of shwfs_im_from_jacobian_noalloc: a file-scope device array, one copy per GPU, held until
// the process ends.  Host code never passes `scratch` itself to a kernel (in host code the name is
// a stand-in, not a device address): the device address is looked up once per device with
// hipGetSymbolAddress while that device is current, and that pointer is used everywhere.
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxScratchDevices] = {};

int scratch_of_current_device(unsigned char** out) {
    int dev = -1;
    SHWFS_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) return set_error(SHWFS_ERR_ARG, "device index outside the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        SHWFS_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return set_error(SHWFS_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return SHWFS_OK;
}

// Launch of the any-size VJP (sh_general_vjp.h) for a plan: the any-size path's own plans, LGS
// plans, and NGS plans whose (np_sub, spec_size) has no specialised VJP kernel.  The first call
// allocates the per-block scratch (at most 2 GiB in total) and, for plans without the any-size
// forward, the twiddle table.
int launch_general_vjp(shwfs_plan* p, const VjpArgs& a, int n_frames, hipStream_t st) {
    const shwfs_geom& g = p->g;
    const int M = g.spec_size;      // a power of two in [16, 1024] (checked by the caller)
    if (!p->d_tw) {
        std::vector<float2> t = twiddle_table(M);
        SHWFS_HIP_CHECK(hipMalloc((void**)&p->d_tw, t.size() * sizeof(float2)));
        SHWFS_HIP_CHECK(hipMemcpy(p->d_tw, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice));
    }
    if (!p->d_vjp_slab) {
        const int64_t sf = gen_vjp_slab(g.np_sub, M, g.subap_npx, g.lgs != 0, p->terms).total;
        const int64_t cap = ((int64_t)2 << 30) / (sf * (int64_t)sizeof(float));
        const int64_t need = (int64_t)g.max_frames * p->n_sub;        // items of the largest call
        const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(1024, cap), need));
        SHWFS_HIP_CHECK(hipMalloc((void**)&p->d_vjp_slab, (size_t)blocks * sf * sizeof(float)));
        p->vjp_slab_floats = sf;
        p->vjp_blocks = blocks;
    }
    GenVjpArgs q{};
    q.tw = p->d_tw;
    q.slab = p->d_vjp_slab;
    q.slab_floats = p->vjp_slab_floats;
    q.M = M;
    int lg = 0;
    while ((1 << lg) < M) ++lg;
    q.log2M = lg;
    q.np = g.np_sub;
    q.log2ch = gen_chunk_log2(M);
    q.ch = 1 << q.log2ch;
    q.terms = p->terms;
    q.term_stride = (int64_t)g.subap_npx * (g.lgs ? 2LL * p->L : (int64_t)M);
    q.kspec = g.lgs ? reinterpret_cast<const float2*>(p->d_kspec) : nullptr;
    VjpArgs b = a;
    if (g.lgs && !p->general) b.roww = p->d_roww_nat;
    const int64_t items = (int64_t)n_frames * p->n_sub;
    const int blocks = (int)std::min<int64_t>(items, p->vjp_blocks);
    if (gen_cap(M) == kGenLdsBigF2)
        hipLaunchKernelGGL(sh_general_vjp_kernel<kGenLdsBigF2>, dim3(blocks), dim3(kGenThreads), 0, st, b, q, n_frames);
    else
        hipLaunchKernelGGL(sh_general_vjp_kernel<kGenLdsF2>, dim3(blocks), dim3(kGenThreads), 0, st, b, q, n_frames);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

// SH VJP of any plan (NGS and LGS, every kernel path).  Non-Parseval plans first run the forward
// into scratch (the image and its unscaled sum give the per-frame scale and cotangent offset,
// vjp_scale_kernel); resampled plans run the VJP on the sensor (resampled) grid and, unless
// `sensor_grid`, scatter it back to the input grid through the transpose of the resampling.  First
// calls allocate the scratch these need (plan-owned, kept).
int vjp_impl(shwfs_plan* p, int n_frames, const float* A, int64_t a_stride, const float* phase, int64_t ph_stride,
             double S0, const float* grad_ccd, int64_t gccd_stride, float* grad_out, int64_t gout_stride,
             bool sensor_grid, hipStream_t st) {
    if (!p || !A || !phase || !grad_ccd || !grad_out) return set_error(SHWFS_ERR_ARG, "null argument");
    const shwfs_geom& g = p->g;
    if (n_frames <= 0 || n_frames > g.max_frames) return set_error(SHWFS_ERR_ARG, "n_frames outside [1, max_frames]");
    const int64_t ef_px = (int64_t)g.ef_size * g.ef_size;                      // sensor grid
    const int64_t in_px = p->ef_in ? (int64_t)p->ef_in * p->ef_in : ef_px;     // input grid
    const int64_t out_px = (sensor_grid || !p->ef_in) ? ef_px : in_px;
    const int64_t ccd_px = (int64_t)g.n_lenses * g.subap_npx * g.n_lenses * g.subap_npx;
    if (n_frames > 1 && ((a_stride != 0 && a_stride < in_px) || (ph_stride != 0 && ph_stride < in_px) ||
                         gccd_stride < ccd_px || gout_stride < out_px))
        return set_error(SHWFS_ERR_ARG, "frame stride too small");
    int ts = 0;
    VjpLauncher L = (g.lgs || p->general) ? nullptr : pick_vjp(g.np_sub, g.spec_size, g.subap_npx, &ts);
    if (!L) {
        const int M = g.spec_size;
        if ((M & (M - 1)) != 0 || M < 16 || M > 1024)
            return set_error(SHWFS_ERR_UNSUPPORTED, "VJP: spectrum grid must be a power of two in [16, 1024]");
        if (g.lgs && !p->kernels_set) return set_error(SHWFS_ERR_ARG, "LGS plan: kernel spectra not set (shwfs_set_kernels)");
    }
    DeviceGuard dg(p->device);
    const int amp_frames = (a_stride == 0) ? 1 : n_frames;
    const double2* cg = nullptr;
    bool cached;
    if (!g.parseval) {
        if (!p->d_vjp_ccd) SHWFS_HIP_CHECK(hipMalloc((void**)&p->d_vjp_ccd, (size_t)g.max_frames * ccd_px * sizeof(float)));
        if (!p->d_vjp_cg) SHWFS_HIP_CHECK(hipMalloc((void**)&p->d_vjp_cg, (size_t)g.max_frames * sizeof(double2)));
        // the forward's own choice of statistics buffers (registered pupil or not)
        cached = A == p->static_A && a_stride == p->static_a_stride && S0 == p->static_S0 &&
                 amp_frames <= p->static_frames;
        int rc = sh_forward_impl(p, n_frames, A, a_stride, phase, ph_stride, S0, p->d_vjp_ccd, ccd_px, nullptr, st);
        if (rc) return rc;
        hipLaunchKernelGGL(vjp_scale_kernel, dim3(n_frames), dim3(1024), 0, st, p->d_bpart, p->n_parts,
                           cached ? p->d_amp_st : p->d_amp, (a_stride == 0) ? 0 : 2, S0, g.pixel_area,
                           p->d_vjp_ccd, ccd_px, grad_ccd, gccd_stride, ccd_px, p->d_vjp_cg);
        SHWFS_LAUNCH_CHECK();
        cg = p->d_vjp_cg;
    } else {
        // sum(A), sum(A^2) do not depend on S0: the registered pupil's statistics serve any S0
        cached = A == p->static_A && a_stride == p->static_a_stride && amp_frames <= p->static_frames;
        if (!cached) {
            int rc = pupil_stats(p, A, a_stride, amp_frames, S0, p->d_amp, p->d_scale, p->d_A_rs, st);
            if (rc) return rc;
        }
        if (p->ef_in) {
            int rc = launch_resample(p, phase, ph_stride, p->d_ext, p->d_ph_rs, n_frames, st);
            if (rc) return rc;
        }
    }
    const float* As = A;
    const float* phs = phase;
    int64_t as = a_stride, ps = ph_stride;
    if (p->ef_in) {      // the sensor-grid field of the forward (resampled amplitude and phase)
        As = cached ? p->d_A_rs_st : p->d_A_rs;
        as = (a_stride == 0) ? 0 : ef_px;
        phs = p->d_ph_rs;
        ps = ef_px;
    }
    float* gs = grad_out;
    int64_t gss = gout_stride;
    if (p->ef_in && !sensor_grid) {
        if (!p->d_gph_rs) SHWFS_HIP_CHECK(hipMalloc((void**)&p->d_gph_rs, (size_t)g.max_frames * ef_px * sizeof(float)));
        gs = p->d_gph_rs;
        gss = ef_px;
    }
    VjpArgs a{};
    a.A = As; a.a_stride = as; a.phase = phs; a.ph_stride = ps;
    a.gccd = grad_ccd; a.gccd_stride = gccd_stride; a.gphase = gs; a.gph_stride = gss;
    a.roww = p->d_roww; a.colw = p->d_colw; a.amp = cached ? p->d_amp_st : p->d_amp;
    a.amp_fstep = (a_stride == 0) ? 0 : 2;
    a.n_lenses = g.n_lenses; a.ef_size = g.ef_size; a.npx = g.subap_npx; a.n_sub = p->n_sub;
    a.k = g.k_rad_per_nm; a.tilt = g.tilt_rad; a.a0 = g.tilt_a0; a.S0 = S0; a.pixel_area = g.pixel_area;
    a.norm_const = g.norm_const;
    a.cg = cg;
    // the subaperture tiles cover the sensor grid (plans have ef_size = n_lenses * np_sub)
    int rc = L ? L(a, (p->n_sub + ts - 1) / ts, n_frames, g.subap_npx, st) : launch_general_vjp(p, a, n_frames, st);
    if (rc) return rc;
    if (p->ef_in && !sensor_grid) {
        const int64_t zb = std::min<int64_t>((in_px + 255) / 256, 2048);
        hipLaunchKernelGGL(zero_frames_kernel, dim3((unsigned)zb, n_frames), dim3(256), 0, st, grad_out, gout_stride,
                           in_px);
        SHWFS_LAUNCH_CHECK();
        const int tiles_x = (g.ef_size + kRvjTX - 1) / kRvjTX, tiles_y = (g.ef_size + kRvjTY - 1) / kRvjTY;
        hipLaunchKernelGGL(resample_vjp_kernel, dim3((unsigned)(tiles_x * tiles_y), n_frames), dim3(256), 0, st, gs,
                           gss, p->d_ext, grad_out, gout_stride, p->ef_in, g.ef_size, p->rmap, tiles_x);
        SHWFS_LAUNCH_CHECK();
    }
    return SHWFS_OK;
}

}  // namespace

extern "C" {

int shwfs_vjp(shwfs_plan* p, int n_frames, const float* A, int64_t a_stride, const float* phase,
              int64_t ph_stride, double S0, const float* grad_ccd, int64_t gccd_stride, float* grad_phase,
              int64_t gph_stride, void* stream) {
    return vjp_impl(p, n_frames, A, a_stride, phase, ph_stride, S0, grad_ccd, gccd_stride, grad_phase, gph_stride,
                    false, reinterpret_cast<hipStream_t>(stream));
}

int shwfs_vjp_sensor_grid(shwfs_plan* p, int n_frames, const float* A, int64_t a_stride, const float* phase,
                          int64_t ph_stride, double S0, const float* grad_ccd, int64_t gccd_stride,
                          float* grad_phase_sensor, int64_t gph_stride, void* stream) {
    return vjp_impl(p, n_frames, A, a_stride, phase, ph_stride, S0, grad_ccd, gccd_stride, grad_phase_sensor,
                    gph_stride, true, reinterpret_cast<hipStream_t>(stream));
}

int shwfs_resample(shwfs_plan* p, int n_frames, const float* in, int64_t in_stride, float* out, int64_t out_stride,
                   void* stream) {
    if (!p || !in || !out) return set_error(SHWFS_ERR_ARG, "null argument");
    if (!p->ef_in) return set_error(SHWFS_ERR_ARG, "plan has no field resampling (shwfs_plan_set_resampling)");
    if (n_frames <= 0 || n_frames > 65535) return set_error(SHWFS_ERR_ARG, "n_frames outside [1, 65535]");
    const int64_t in_px = (int64_t)p->ef_in * p->ef_in, out_px = (int64_t)p->g.ef_size * p->g.ef_size;
    if (n_frames > 1 && ((in_stride != 0 && in_stride < in_px) || out_stride < out_px))
        return set_error(SHWFS_ERR_ARG, "frame stride too small");
    DeviceGuard dg(p->device);
    const int64_t blocks = std::min<int64_t>((out_px + 255) / 256, 2048);
    hipLaunchKernelGGL(resample_kernel, dim3((unsigned)blocks, n_frames), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), in, in_stride, p->d_ext, out, out_stride, p->ef_in,
                       p->g.ef_size, p->rmap);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
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
                           float* im, void* work, void* stream) {
    int rc = im_check_args(n_lenses, np_sub, ef_size, jx, jy, n_modes, modes, mode_stride, slot, n_valid, im);
    if (rc) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (!im_strip_kernels(np_sub, ef_size))
        return im_launch_any(n_lenses, np_sub, ef_size, jx, jy, n_modes, modes, mode_stride, slot, n_valid, scale, im,
                             st);
    if (!im_v4_eligible(np_sub, ef_size, mode_stride, modes, jx, jy))
        return im_launch_scalar(n_lenses, np_sub, ef_size, jx, jy, n_modes, modes, mode_stride, slot, n_valid, scale,
                                im, st);
    const int64_t nn = (int64_t)n_lenses * n_lenses;
    float* gridbuf = static_cast<float*>(work);
    int32_t* inv = work ? reinterpret_cast<int32_t*>(gridbuf + (int64_t)n_modes * 2 * nn) : nullptr;
    return im_launch_v4(n_lenses, np_sub, ef_size, jx, jy, n_modes, modes, mode_stride, slot, n_valid, scale, im,
                        gridbuf, inv, n_modes, st);
}

int shwfs_im_from_jacobian_noalloc(int n_lenses, int np_sub, int ef_size, const float* jx, const float* jy,
                                   int n_modes, const float* modes, int64_t mode_stride, const int32_t* slot,
                                   int n_valid, float scale, float* im, void* stream) {
    int rc = im_check_args(n_lenses, np_sub, ef_size, jx, jy, n_modes, modes, mode_stride, slot, n_valid, im);
    if (rc) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (!im_strip_kernels(np_sub, ef_size))                              // these paths need no scratch
        return im_launch_any(n_lenses, np_sub, ef_size, jx, jy, n_modes, modes, mode_stride, slot, n_valid, scale, im,
                             st);
    if (!im_v4_eligible(np_sub, ef_size, mode_stride, modes, jx, jy))
        return im_launch_scalar(n_lenses, np_sub, ef_size, jx, jy, n_modes, modes, mode_stride, slot, n_valid, scale,
                                im, st);
    const int64_t nn = (int64_t)n_lenses * n_lenses;
    const size_t inv_bytes = ((size_t)n_valid * sizeof(int32_t) + 255) / 256 * 256;
    const size_t mode_bytes = (size_t)2 * nn * sizeof(float);        // grid-order results of one mode
    if (inv_bytes + mode_bytes > kScratchBytes)
        return set_error(SHWFS_ERR_ARG, "shwfs_im_from_jacobian_noalloc: n_valid and n_lenses^2 need more than the "
                                        "16 MiB scratch for one mode (n_valid * 4 + 8 n_lenses^2 bytes)");
    unsigned char* base = nullptr;
    rc = scratch_of_current_device(&base);
    if (rc) return rc;
    int32_t* inv = reinterpret_cast<int32_t*>(base);
    float* gridbuf = reinterpret_cast<float*>(base + inv_bytes);
    const int chunk = (int)std::min<size_t>((size_t)n_modes, (kScratchBytes - inv_bytes) / mode_bytes);
    return im_launch_v4(n_lenses, np_sub, ef_size, jx, jy, n_modes, modes, mode_stride, slot, n_valid, scale, im,
                        gridbuf, inv, chunk, st);
}

}  // extern "C"

struct shwfs_batch {
    int device = 0;
    int n_sensors = 0;
    int n_items = 0;
    int grid = 0;
    int n_ngs = 0, n_lgs = 0;         // items of each kind: d_items[0, n_lgs) LGS, then NGS
    NgsArgs* d_args = nullptr;
    BatchItem* d_items = nullptr;
    double* d_amp = nullptr;          // [n_sensors][2]
    float* d_scale = nullptr;         // [n_sensors]
    double* d_amp_part = nullptr;
    unsigned* d_counter = nullptr;
    unsigned long long* d_ts = nullptr;   // [1 + n_sensors]
    int32_t* d_order = nullptr;       // lit-first subaperture order of each sensor's pupil, concatenated
    uint8_t* d_lit = nullptr;         // lit flags (scratch of the order computation)
    struct Resample {                 // resampled sensor: its phase frames resampled before the SH launches
        const shwfs_plan* plan;
        const float* phase;           // input-grid frames, ef_in^2 apart
        int n_frames;
        float* ph_rs;                 // sensor-grid frames, ef_size^2 apart
    };
    struct Rescale {                  // non-Parseval sensor: its images rescaled after the SH launches
        float* ccd;
        int64_t ccd_px;
        double* bpart;
        int n_parts;
        const double* amp;
        double S0, pixel_area;
        int n_frames;
    };
    struct Direct {                   // sensor outside the persistent bodies: its plan's own launch
        shwfs_plan* plan;
        int sensor;
        const float* A;
        const float* phase;
        float* ccd;
        double S0;
        int n_frames;
    };
    struct Slopes {                   // sensor whose ShSlopec runs in the SH launch's epilogue
        const shslopec_plan* c;
        double* tpart;                // [n_frames][n_parts][2] per-wave partial totals
        int n_parts;
        float* slopes;
        int64_t slopes_stride;
        double* flux;
        double* totals;
        int n_frames;
    };
    std::vector<Resample> resample;
    std::vector<Rescale> rescale;
    std::vector<Direct> direct;
    std::vector<Slopes> slopes;
    std::vector<NgsArgs> h_args;      // host copy of d_args
    std::vector<shwfs_plan*> plans;
    std::vector<int> kinds, frames;
    std::vector<void*> bufs;          // device buffers of the above (resampled fields, partial sums)
    std::vector<BatchItem> h_items;   // the work list in the requested order (LGS items, then NGS)
    int grid_req = 0;                 // blocks per launch requested (0: one per item)
    struct Launch {                   // one persistent launch: items [offset, offset + count) of d_items
        int kind, variant, offset, count;
    };
    std::vector<Launch> launches;
    CogFinalizeRec* d_fin = nullptr;  // fused-CoG sensors' finalize records, one launch for all
    int2* d_fin_map = nullptr;        // [fin_rows]: (record, frame)
    int fin_rows = 0, fin_bx = 0;
};

namespace {
// Groups the work list by body variant (each variant is a kernel of its own, over a contiguous
// range of d_items; the requested order holds within a variant) and orders the launches: LGS
// before NGS, within a kind the variants with fewer items first, so that the sensors of a
// real-time step complete before a large calibration batch.  Uploads d_items (synchronous).
int batch_layout(shwfs_batch* b) {
    b->launches.clear();
    std::vector<BatchItem> sorted;
    sorted.reserve(b->h_items.size());
    for (int kind = 1; kind >= 0; --kind) {
        std::vector<std::vector<BatchItem>> per(kBatchVariants[kind]);
        for (const auto& it : b->h_items)
            if (b->kinds[it.sensor] == kind) per[batch_variant(b->h_args[it.sensor], kind)].push_back(it);
        std::vector<int> vs;
        for (int v = 0; v < kBatchVariants[kind]; ++v)
            if (!per[v].empty()) vs.push_back(v);
        std::stable_sort(vs.begin(), vs.end(), [&](int x, int y) { return per[x].size() < per[y].size(); });
        for (const int v : vs) {
            b->launches.push_back({kind, v, (int)sorted.size(), (int)per[v].size()});
            sorted.insert(sorted.end(), per[v].begin(), per[v].end());
        }
    }
    const hipError_t e = hipMemcpy(b->d_items, sorted.data(), sorted.size() * sizeof(BatchItem), hipMemcpyHostToDevice);
    if (e != hipSuccess) return set_error(SHWFS_ERR_HIP, std::string("batch items: ") + hipGetErrorString(e));
    return SHWFS_OK;
}
}  // namespace

extern "C" {

int shwfs_batch_destroy(shwfs_batch* b) {
    if (!b) return SHWFS_OK;
    DeviceGuard dg(b->device);
    (void)hipFree(b->d_args);
    (void)hipFree(b->d_items);
    (void)hipFree(b->d_fin);
    (void)hipFree(b->d_fin_map);
    (void)hipFree(b->d_amp);
    (void)hipFree(b->d_scale);
    (void)hipFree(b->d_amp_part);
    (void)hipFree(b->d_counter);
    (void)hipFree(b->d_ts);
    (void)hipFree(b->d_order);
    (void)hipFree(b->d_lit);
    for (void* q : b->bufs) (void)hipFree(q);
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
    std::vector<int> cost(n_sensors), kinds(n_sensors), tiles_of(n_sensors);
    for (int s = 0; s < n_sensors; ++s) {
        const shwfs_plan* p = plans[s];
        if (!p || !A[s] || !phase[s] || !ccd[s]) return set_error(SHWFS_ERR_ARG, "null sensor entry");
        if (p->device != device) return set_error(SHWFS_ERR_ARG, "plan on another device");
        const shwfs_geom& g = p->g;
        int kind;
        if (!g.lgs && p->mfma && g.np_sub == 8 && g.spec_size == 16) kind = 0;
        else if (g.lgs && !p->general && g.np_sub == 16 && g.spec_size == 32) kind = 1;
        else kind = 2;       // any other plan: its own launch in the batch's sequence
        if (g.lgs && !p->kernels_set) return set_error(SHWFS_ERR_ARG, "LGS plan without kernels");
        if (n_frames[s] <= 0 || n_frames[s] > 65535) return set_error(SHWFS_ERR_ARG, "frame count outside [1, 65535]");
        if (kind == 2) {
            if (n_frames[s] > g.max_frames)
                return set_error(SHWFS_ERR_ARG, "frame count above the plan's max_frames (sensor run with its own launch)");
            kinds[s] = 2;
            cost[s] = 0;
            tiles_of[s] = 0;
            args[s] = NgsArgs{};
            continue;
        }
        const int ts = (kind == 0) ? 4 * NgsMfmaShape<8>::SW : 8;
        const int tiles = (p->n_sub + ts - 1) / ts;
        tiles_of[s] = tiles;
        cost[s] = kind == 0 ? 1 : 4;
        NgsArgs& a = args[s];
        a = NgsArgs{};
        a.A = A[s]; a.a_stride = 0; a.phase = phase[s]; a.ph_stride = (int64_t)g.ef_size * g.ef_size;
        a.ccd = ccd[s];
        a.ccd_stride = (int64_t)g.n_lenses * g.subap_npx * g.n_lenses * g.subap_npx;
        a.roww = (kind == 0) ? p->d_rowlag : p->d_roww; a.colw = p->d_colw; a.amp_fstep = 0; a.bpart = nullptr;
        a.kspec = reinterpret_cast<const float2*>(p->d_kspec);
        a.n_lenses = g.n_lenses; a.ef_size = g.ef_size; a.npx = g.subap_npx; a.npx_pad = p->npx_pad;
        a.n_sub = p->n_sub; a.n_tiles = tiles; a.parseval = g.parseval ? 1 : 0;
        a.n_parts = (kind == 0) ? tiles * NgsMfmaShape<8>::WPB : tiles;
        a.store_vec = aligned16(ccd[s]) && ((int64_t)g.n_lenses * g.subap_npx) % 4 == 0;
        a.k = g.k_rad_per_nm; a.tilt = g.tilt_rad; a.a0 = g.tilt_a0; a.S0 = S0[s];
        a.pixel_area = g.pixel_area; a.norm_const = g.norm_const;
        kinds[s] = kind;
        const int fstep = (kind == 0) ? kBatchNgsFpb : 1;
        for (int f = 0; f < n_frames[s]; f += fstep)
            for (int t = 0; t < tiles; ++t) items.push_back(BatchItem{s, f, t, std::min(fstep, n_frames[s] - f)});
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
    // LGS items first (their own launch), then NGS items; the order above holds within each kind
    std::stable_partition(items.begin(), items.end(), [&](const BatchItem& x) { return kinds[x.sensor] == 1; });
    DeviceGuard dg(device);
    auto* b = new shwfs_batch();
    b->device = device;
    b->n_sensors = n_sensors;
    b->n_items = (int)items.size();
    b->n_lgs = (int)std::count_if(items.begin(), items.end(), [&](const BatchItem& x) { return kinds[x.sensor] == 1; });
    b->n_ngs = b->n_items - b->n_lgs;
    b->grid = grid_blocks > 0 ? std::min(grid_blocks, b->n_items) : b->n_items;
    auto cleanup = [&](int code) { shwfs_batch_destroy(b); return code; };
    hipError_t e;
    const int amp_blocks = 256;
    if ((e = hipMalloc((void**)&b->d_args, n_sensors * sizeof(NgsArgs))) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_items, items.size() * sizeof(BatchItem))) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_amp, n_sensors * 2 * sizeof(double))) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_scale, n_sensors * sizeof(float))) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_amp_part, n_sensors * amp_blocks * 2 * sizeof(double))) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_counter, n_sensors * sizeof(unsigned))) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_ts, (n_sensors + 1) * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipMemset(b->d_counter, 0, n_sensors * sizeof(unsigned))) != hipSuccess)
        return cleanup(set_error(SHWFS_ERR_HIP, std::string("batch alloc: ") + hipGetErrorString(e)));
    size_t n_order = 0;
    for (int s = 0; s < n_sensors; ++s) n_order += (size_t)plans[s]->n_sub;
    if ((e = hipMalloc((void**)&b->d_order, n_order * sizeof(int32_t))) != hipSuccess ||
        (e = hipMalloc((void**)&b->d_lit, n_order)) != hipSuccess)
        return cleanup(set_error(SHWFS_ERR_HIP, std::string("batch alloc: ") + hipGetErrorString(e)));
    auto alloc = [&](size_t bytes, void** q) {
        *q = nullptr;
        e = hipMalloc(q, bytes);
        if (e == hipSuccess) b->bufs.push_back(*q);
        return e == hipSuccess;
    };
    // the pupils are static for the batch: amplitude statistics (resampled sensors: on the
    // resampled amplitude, as pupil_stats) and the lit-first subaperture order (Parseval sensors:
    // their images do not depend on the grouping) once, here
    size_t order_off = 0;
    for (int s = 0; s < n_sensors; ++s) {
        const shwfs_plan* p = plans[s];
        if (kinds[s] == 2) {
            b->direct.push_back({plans[s], s, A[s], phase[s], ccd[s], S0[s], n_frames[s]});
            order_off += (size_t)p->n_sub;
            continue;
        }
        const shwfs_geom& gs = p->g;
        const int64_t px = (int64_t)gs.ef_size * gs.ef_size;
        const float* A_grid = A[s];                   // amplitude on the sensor grid
        const float* B = nullptr;
        int64_t n_in = px;
        if (p->ef_in) {
            void *a_rs = nullptr, *ph_rs = nullptr;
            if (!alloc((size_t)px * sizeof(float), &a_rs) ||
                !alloc((size_t)n_frames[s] * px * sizeof(float), &ph_rs))
                return cleanup(set_error(SHWFS_ERR_HIP, std::string("batch alloc: ") + hipGetErrorString(e)));
            const int rc = launch_resample(p, A[s], 0, nullptr, static_cast<float*>(a_rs), 1, 0);
            if (rc) return cleanup(rc);
            A_grid = static_cast<const float*>(a_rs);
            B = A_grid;
            n_in = (int64_t)p->ef_in * p->ef_in;
            args[s].A = A_grid;
            args[s].phase = static_cast<const float*>(ph_rs);
            b->resample.push_back({p, phase[s], n_frames[s], static_cast<float*>(ph_rs)});
        }
        args[s].vec = (gs.ef_size % 4 == 0) && aligned16(args[s].A) && aligned16(args[s].phase);
        int32_t* ord = b->d_order + order_off;
        const int rco = launch_subap_order(A_grid, gs.ef_size, gs.n_lenses, gs.np_sub, p->n_sub,
                                           b->d_lit + order_off, ord, 0);
        if (rco) return cleanup(rco);
        order_off += (size_t)p->n_sub;
        args[s].order = args[s].parseval ? ord : nullptr;
        hipLaunchKernelGGL(amp_stats_kernel, dim3(amp_blocks, 1), dim3(256), 0, 0, A[s], (int64_t)0, n_in,
                           b->d_amp_part + (size_t)s * amp_blocks * 2, b->d_counter + s, b->d_amp + 2 * s,
                           S0[s] * gs.pixel_area, gs.norm_const, gs.parseval ? b->d_scale + s : nullptr, B,
                           (int64_t)0, B ? px : (int64_t)0);
        args[s].amp = b->d_amp + 2 * s;
        args[s].fscale = b->d_scale + s;
        args[s].scale_fstep = 0;
        if (!gs.parseval) {
            void* bp = nullptr;
            if (!alloc((size_t)n_frames[s] * args[s].n_parts * sizeof(double), &bp))
                return cleanup(set_error(SHWFS_ERR_HIP, std::string("batch alloc: ") + hipGetErrorString(e)));
            args[s].bpart = static_cast<double*>(bp);
            b->rescale.push_back({ccd[s], args[s].ccd_stride, static_cast<double*>(bp), args[s].n_parts,
                                  b->d_amp + 2 * s, S0[s], gs.pixel_area, n_frames[s]});
        }
    }
    if ((e = hipGetLastError()) != hipSuccess || (e = hipDeviceSynchronize()) != hipSuccess ||
        (e = hipMemcpy(b->d_args, args.data(), n_sensors * sizeof(NgsArgs), hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(set_error(SHWFS_ERR_HIP, std::string("batch setup: ") + hipGetErrorString(e)));
    b->h_args = args;
    b->plans.assign(plans, plans + n_sensors);
    b->kinds = kinds;
    b->frames.assign(n_frames, n_frames + n_sensors);
    b->h_items = items;
    b->grid_req = grid_blocks > 0 ? grid_blocks : 0;
    if (const int rc = batch_layout(b)) return cleanup(rc);
    *out = b;
    return SHWFS_OK;
}

int shwfs_batch_set_slopes(shwfs_batch* b, int sensor, float thr, int thr_mode, float* slopes,
                           int64_t slopes_stride, double* flux, double* totals) {
    if (!b || !slopes || !flux || !totals) return set_error(SHWFS_ERR_ARG, "null argument");
    if (sensor < 0 || sensor >= b->n_sensors) return set_error(SHWFS_ERR_ARG, "sensor index out of range");
    if (thr_mode != 0 && thr_mode != 1) return set_error(SHWFS_ERR_ARG, "thr_mode must be 0 or 1");
    shwfs_plan* p = b->plans[sensor];
    const shslopec_plan* c = p->cog;
    NgsArgs& a = b->h_args[sensor];
    if (!c) return set_error(SHWFS_ERR_ARG, "no ShSlopec plan bound to the sensor's plan (shwfs_plan_bind_slopec)");
    const bool ngs_ok = b->kinds[sensor] == 0 && a.store_vec && (a.npx == 8 || a.npx == 16);
    const bool lgs_ok = b->kinds[sensor] == 1 && a.npx <= 16;
    if (!a.parseval || !(ngs_ok || lgs_ok))
        return set_error(SHWFS_ERR_UNSUPPORTED, "fused CoG in a batch needs Parseval normalisation and an NGS MFMA "
                                                "np 8 sensor with 8 or 16 px per subaperture and 16-byte aligned CCD "
                                                "rows, or an LGS sensor of the batch's LGS body");
    if (a.cog) return set_error(SHWFS_ERR_ARG, "the sensor's ShSlopec already runs in the batch");
    for (const auto& r : b->slopes)
        if (r.c == c && r.tpart) return set_error(SHWFS_ERR_ARG, "this ShSlopec plan already serves a sensor of the batch");
    const int nf = b->frames[sensor];
    if (nf > c->max_frames) return set_error(SHWFS_ERR_ARG, "frame count above the ShSlopec plan's max_frames");
    if (nf > 1 && slopes_stride < 2 * (int64_t)c->n_sub) return set_error(SHWFS_ERR_ARG, "slopes frame stride too small");
    DeviceGuard dg(b->device);
    if (b->kinds[sensor] == 1) a.n_parts = a.n_tiles * 8;   // LGS: one partial per tile position
    void* tp = nullptr;
    hipError_t e = hipMalloc(&tp, (size_t)nf * a.n_parts * 2 * sizeof(double));
    if (e != hipSuccess) return set_error(SHWFS_ERR_HIP, std::string("batch alloc: ") + hipGetErrorString(e));
    b->bufs.push_back(tp);
    a.cog = thr_mode + 1;
    a.thr = thr;
    a.slot = p->d_slot;
    a.mom = c->d_mom;
    a.cogw = p->d_cogw;
    a.tpart = static_cast<double*>(tp);
    a.n_valid = c->n_sub;
    SHWFS_HIP_CHECK(hipMemcpy(b->d_args + sensor, &a, sizeof(NgsArgs), hipMemcpyHostToDevice));
    b->slopes.push_back({c, static_cast<double*>(tp), a.n_parts, slopes, slopes_stride, flux, totals, nf});
    // the finalize records of every fused sensor, for one launch after the SH launches
    std::vector<CogFinalizeRec> recs;
    std::vector<int2> map;
    int max_sub = 0;
    for (const auto& r : b->slopes) {
        for (int f = 0; f < r.n_frames; ++f) map.push_back(make_int2((int)recs.size(), f));
        recs.push_back({r.c->d_mom, r.tpart, r.n_parts, r.c->n_sub, r.slopes, r.slopes_stride, r.flux, r.totals});
        max_sub = std::max(max_sub, r.c->n_sub);
    }
    (void)hipFree(b->d_fin); b->d_fin = nullptr;
    (void)hipFree(b->d_fin_map); b->d_fin_map = nullptr;
    SHWFS_HIP_CHECK(hipMalloc((void**)&b->d_fin, recs.size() * sizeof(CogFinalizeRec)));
    SHWFS_HIP_CHECK(hipMalloc((void**)&b->d_fin_map, map.size() * sizeof(int2)));
    SHWFS_HIP_CHECK(hipMemcpy(b->d_fin, recs.data(), recs.size() * sizeof(CogFinalizeRec), hipMemcpyHostToDevice));
    SHWFS_HIP_CHECK(hipMemcpy(b->d_fin_map, map.data(), map.size() * sizeof(int2), hipMemcpyHostToDevice));
    b->fin_rows = (int)map.size();
    b->fin_bx = (max_sub + 1023) / 1024;
    return batch_layout(b);        // the sensor's items now run the fused-CoG variant of its body
}

int shwfs_batch_forward(shwfs_batch* b, void* stream) {
    if (!b) return set_error(SHWFS_ERR_ARG, "null batch");
    DeviceGuard dg(b->device);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(batch_stamp_kernel, dim3(1), dim3(64), 0, st, b->d_ts, b->n_sensors);
    SHWFS_LAUNCH_CHECK();
    for (const auto& r : b->resample) {     // the resampled sensors' phases onto their sensor grids
        const int rc = launch_resample(r.plan, r.phase, (int64_t)r.plan->ef_in * r.plan->ef_in, r.plan->d_ext, r.ph_rs,
                                       r.n_frames, st);
        if (rc) return rc;
    }
    for (const auto& L : b->launches) {     // the persistent SH launches (batch_layout)
        const BatchKernel k = L.kind == 1 ? batch_kernel_of<1>(L.variant, std::make_integer_sequence<int, kBatchVariants[1]>{})
                                          : batch_kernel_of<0>(L.variant, std::make_integer_sequence<int, kBatchVariants[0]>{});
        const int grid = b->grid_req > 0 ? std::min(b->grid_req, L.count) : L.count;
        hipLaunchKernelGGL(k, dim3(grid), dim3(256), L.kind == 1 ? kBatchLgsLds : kBatchNgsLds, st, b->d_args,
                           b->d_items + L.offset, L.count, b->d_ts);
        SHWFS_LAUNCH_CHECK();
    }
    if (b->fin_rows > 0) {                  // fused-CoG sensors: slopes, flux and totals from the moments
        hipLaunchKernelGGL(cog_finalize_batch_kernel, dim3(b->fin_bx, std::min(b->fin_rows, 65535)), dim3(256), 0, st,
                           b->d_fin, b->d_fin_map, b->fin_rows);
        SHWFS_LAUNCH_CHECK();
    }
    for (const auto& c : b->rescale) {      // non-Parseval sensors: images normalised by their sums
        hipLaunchKernelGGL(sh_rescale_kernel, dim3(64, c.n_frames), dim3(256), 0, st, c.ccd, c.ccd_px, c.ccd_px,
                           c.bpart, c.n_parts, c.amp, 0, c.S0, c.pixel_area);
        SHWFS_LAUNCH_CHECK();
    }
    for (const auto& d : b->direct) {       // sensors outside the persistent bodies
        const shwfs_geom& g = d.plan->g;
        const int64_t in_px = d.plan->ef_in ? (int64_t)d.plan->ef_in * d.plan->ef_in : (int64_t)g.ef_size * g.ef_size;
        const int64_t ccd_px = (int64_t)g.n_lenses * g.subap_npx * g.n_lenses * g.subap_npx;
        const int rc = sh_forward_impl(d.plan, d.n_frames, d.A, 0, d.phase, in_px, d.S0, d.ccd, ccd_px, nullptr, st);
        if (rc) return rc;
        hipLaunchKernelGGL(batch_sensor_stamp_kernel, dim3(1), dim3(64), 0, st, b->d_ts, d.sensor);
        SHWFS_LAUNCH_CHECK();
    }
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

// ============================================================================ HIP graphs
// Capture of library calls into a HIP graph: the caller begins a capture on its (non-default)
// stream, issues any shwfs_* / shslopec_* / shccd_* calls on that stream, and ends it; replays
// are one hipGraphLaunch.  Thread-local capture mode: only calls from the capturing thread are
// recorded, and other threads may keep allocating.  The library's kernels clear their own
// counters (no memset nodes in a graph).
struct shwfs_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

int shwfs_graph_begin(void* stream) {
    if (!stream) return set_error(SHWFS_ERR_ARG, "graph capture needs a non-default stream");
    SHWFS_HIP_CHECK(hipStreamBeginCapture(reinterpret_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal));
    return SHWFS_OK;
}

int shwfs_graph_end(void* stream, shwfs_graph** out) {
    if (!stream || !out) return set_error(SHWFS_ERR_ARG, "null argument");
    *out = nullptr;
    hipGraph_t graph = nullptr;
    SHWFS_HIP_CHECK(hipStreamEndCapture(reinterpret_cast<hipStream_t>(stream), &graph));
    if (!graph) return set_error(SHWFS_ERR_HIP, "stream capture produced no graph");
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return set_error(SHWFS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    auto* g = new shwfs_graph();
    g->graph = graph;
    g->exec = exec;
    *out = g;
    return SHWFS_OK;
}

int shwfs_graph_launch(shwfs_graph* g, void* stream) {
    if (!g || !g->exec) return set_error(SHWFS_ERR_ARG, "null graph");
    SHWFS_HIP_CHECK(hipGraphLaunch(g->exec, reinterpret_cast<hipStream_t>(stream)));
    return SHWFS_OK;
}

int shwfs_graph_node_count(const shwfs_graph* g, int64_t* n_nodes) {
    if (!g || !g->graph || !n_nodes) return set_error(SHWFS_ERR_ARG, "null argument");
    size_t n = 0;
    SHWFS_HIP_CHECK(hipGraphGetNodes(g->graph, nullptr, &n));
    *n_nodes = (int64_t)n;
    return SHWFS_OK;
}

int shwfs_graph_destroy(shwfs_graph* g) {
    if (!g) return SHWFS_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return SHWFS_OK;
}

}  // extern "C"

extern "C" {

int shwfs_pushpull_fields(int n_modes, const float* phi0, int64_t n_px, const float* modes, int64_t mode_stride,
                          float amp, float* out, int64_t out_stride, void* stream) {
    if (!phi0 || !modes || !out) return set_error(SHWFS_ERR_ARG, "null argument");
    if (n_modes <= 0 || n_modes > 65535 || n_px <= 0) return set_error(SHWFS_ERR_ARG, "n_modes outside [1, 65535] or empty field");
    if ((n_modes > 1 && mode_stride < n_px) || out_stride < n_px)
        return set_error(SHWFS_ERR_ARG, "frame stride too small");
    const int64_t blocks = std::min<int64_t>((n_px + 255) / 256, 1024);
    hipLaunchKernelGGL(pushpull_fields_kernel, dim3((unsigned)blocks, n_modes), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), phi0, modes, mode_stride, n_px, amp, out, out_stride,
                       n_modes);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

int shwfs_pushpull_rows(int n_modes, const float* slopes, int64_t slope_stride, int n_slopes, float amp, float* im,
                        int64_t im_stride, void* stream) {
    if (!slopes || !im) return set_error(SHWFS_ERR_ARG, "null argument");
    if (n_modes <= 0 || n_modes > 65535 || n_slopes <= 0) return set_error(SHWFS_ERR_ARG, "n_modes outside [1, 65535] or no slopes");
    if (slope_stride < n_slopes || (n_modes > 1 && im_stride < n_slopes))
        return set_error(SHWFS_ERR_ARG, "row stride too small");
    if (!(amp != 0.f)) return set_error(SHWFS_ERR_ARG, "push-pull amplitude must be non-zero");
    const int blocks = std::min((n_slopes + 255) / 256, 64);
    hipLaunchKernelGGL(pushpull_rows_kernel, dim3(blocks, n_modes), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       slopes, slope_stride, n_slopes, amp, im, im_stride, n_modes);
    SHWFS_LAUNCH_CHECK();
    return SHWFS_OK;
}

}  // extern "C"
