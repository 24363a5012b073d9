// ccd.hip -- detector post-step of the SH path on gfx950: CCD integration and readout
// (processing_objects/ccd.py:198-280) as two element-wise kernels.
//
// Readout of one pixel (ccd.py:210-258, in the reference's order):
//   x  = b^2 * (sum of the b x b block of the integrated frame)        apply_binning
//   x *= QE                                                            apply_qe
//   x += background + dark                       (either flag)
//   x += Bernoulli(cic_level)                    (cic)
//   x  = Poisson(x)                              (photon; the frame becomes int64)
//   x  = Gamma(shape delta*x, scale emccd) / delta  (excess)
//   x += readout * N(0,1)   -- truncated toward zero while the frame is int64 (photon without
//                              excess), because the reference adds `.astype(ccd_frame.dtype)`
//   x *= pixel gain
//   photon: x = max(round_half_even(x*ADU_gain) + bias, 0) [- bias] / ADU_gain [/ emccd]
//           [- dark] [- background]
// Random numbers come from a counter-based generator (Philox4x32-10, Salmon et al., SC'11),
// keyed by the seed of each noise source and indexed by (pixel, readout counter, draw), so any
// pixel of any readout is reproducible without sequential state.  The streams are not numpy's:
// parity with the reference is distributional.
//
// Poisson sampling: inversion of the CDF (sequential search) for lambda < 10, Hormann's transformed
// rejection with squeeze (PTRS, Insurance: Mathematics and Economics 12 (1993) 39-45) above, in
// FP32 with a cancellation-free log-pmf (the count itself is kept in FP64).  Gamma: Marsaglia & Tsang (ACM TOMS 26 (2000) 363-372).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/shwfs.h"

int shwfs_internal_set_error(int code, const char* msg);

namespace {

#define CCD_HIP_CHECK(expr)                                                                  \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return shwfs_internal_set_error(SHWFS_ERR_HIP, (std::string(#expr) + ": " +       \
                                                            hipGetErrorString(e_)).c_str()); \
    } while (0)

// ---------------------------------------------------------------- Philox4x32-10
struct U4 {
    uint32_t x, y, z, w;
};

// a ^ b ^ c in one instruction (v_bitop3_b32 with the truth table of a three-input xor)
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// one round: two 32x32 -> 64 multiplies (one v_mad_u64_u32 each) and two three-input xors
__device__ __forceinline__ U4 philox(U4 c, uint32_t k0, uint32_t k1) {
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c.x, p1 = (uint64_t)M1 * c.z;
        c = U4{xor3((uint32_t)(p1 >> 32), c.y, k0), (uint32_t)p1, xor3((uint32_t)(p0 >> 32), c.w, k1), (uint32_t)p0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

__global__ void philox_kat_kernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* c = in + 6 * i;
    const U4 r = philox(U4{c[0], c[1], c[2], c[3]}, c[4], c[5]);
    out[4 * i] = r.x;
    out[4 * i + 1] = r.y;
    out[4 * i + 2] = r.z;
    out[4 * i + 3] = r.w;
}

// Stream of uniforms of one pixel for one noise source: draw d -> 4 words per Philox call.
struct Stream {
    uint32_t k0, k1;          // seed
    uint32_t px_lo, px_hi;    // pixel index
    uint32_t ctr;             // readout counter
    uint32_t draw = 0;
    U4 buf{};
    int left = 0;

    __device__ Stream(uint64_t seed, uint64_t px, uint32_t counter, uint32_t first_draw = 0)
        : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), px_lo((uint32_t)px), px_hi((uint32_t)(px >> 32)),
          ctr(counter), draw(first_draw) {}
    __device__ Stream(uint32_t key0, uint32_t key1, uint64_t px, uint32_t counter, uint32_t first_draw)
        : k0(key0), k1(key1), px_lo((uint32_t)px), px_hi((uint32_t)(px >> 32)), ctr(counter), draw(first_draw) {}

    __device__ __forceinline__ uint32_t next_u32() {
        if (left == 0) {
            buf = philox(U4{px_lo, px_hi, ctr, draw++}, k0, k1);
            left = 4;
        }
        --left;
        return left == 3 ? buf.x : left == 2 ? buf.y : left == 1 ? buf.z : buf.w;
    }
    // (0, 1] with 24 bits (single) and 53 bits (double)
    __device__ __forceinline__ float uniform_f() { return ((next_u32() >> 8) + 1u) * (1.0f / 16777216.0f); }
    __device__ __forceinline__ double uniform_d() {
        const uint64_t hi = next_u32() >> 5, lo = next_u32() >> 6;     // 27 + 26 bits
        return ((double)((hi << 26) | lo) + 1.0) * (1.0 / 9007199254740992.0);
    }
    __device__ __forceinline__ float normal() {                          // Box-Muller, one of the pair
        const float u1 = uniform_f(), u2 = uniform_f();
        return sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
    }
};

// ---------------------------------------------------------------- log k!
__constant__ double kLogFact[32] = {
    0.0, 0.0, 0.693147180559945, 1.7917594692280554,
    3.178053830347945, 4.787491742782047, 6.579251212010102, 8.525161361065415,
    10.604602902745249, 12.801827480081467, 15.104412573075514, 17.502307845873887,
    19.987214495661885, 22.55216385312342, 25.191221182738683, 27.89927138384089,
    30.671860106080672, 33.50507345013689, 36.39544520803305, 39.339884187199495,
    42.335616460753485, 45.38013889847691, 48.47118135183522, 51.60667556776438,
    54.78472939811232, 58.00360522298052, 61.26170176100201, 64.55753862700634,
    67.88974313718153, 71.257038967168, 74.65823634883017, 78.0922235533153};

__device__ __forceinline__ float u01(uint32_t w) { return ((w >> 8) + 1u) * (1.0f / 16777216.0f); }   // (0, 1]

// N(0,1) from two uniform words (Box-Muller; hardware log2 and cos of revolutions).
__device__ __forceinline__ float normal_from(uint32_t w0, uint32_t w1) {
    const float r = __builtin_sqrtf(-2.0f * 0.69314718056f * __builtin_amdgcn_logf(u01(w0)));
    return r * __builtin_amdgcn_cosf(u01(w1));
}

__device__ __forceinline__ float fast_log(float x) { return 0.69314718056f * __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// log of the Poisson pmf at k >= 32 without cancellation: with k = lam (1 + x) and Stirling's
// series for log k!,  log p = -lam phi(x) - log(2 pi k)/2 - 1/(12k) + 1/(360k^3),
// phi(x) = (1 + x) log(1 + x) - x = sum_{n>=2} (-x)^n / (n (n-1))  (series below |x| = 0.1).
__device__ __forceinline__ float log_pmf_large(double k, float lam, float inv_lam) {
    const float x = (float)(k - (double)lam) * inv_lam;
    float phi;
    if (fabsf(x) < 0.1f)
        phi = x * x * (0.5f - x * (1.0f / 6 - x * (1.0f / 12 - x * (1.0f / 20 - x * (1.0f / 30 - x * (1.0f / 42 - x * (1.0f / 56)))))));
    else
        phi = (1.0f + x) * fast_log(1.0f + x) - x;
    const float kf = (float)k, r = fast_rcp(kf);
    return -lam * phi - 0.5f * fast_log(6.283185307f * kf) - r * (1.0f / 12 - r * r * (1.0f / 360));
}

__device__ double gamma_sample(Stream& s, double shape, double scale) {
    if (!(shape > 0.0)) return 0.0;
    double boost = 1.0;
    if (shape < 1.0) {                                    // G(a) = G(a + 1) U^(1/a)
        boost = pow(s.uniform_d(), 1.0 / shape);
        shape += 1.0;
    }
    const double d = shape - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
    for (int it = 0; it < 256; ++it) {
        double x = 0.0, v = 1.0;
        for (int t = 0; t < 64; ++t) {                    // P(v <= 0) < 1% per draw
            x = (double)s.normal();
            v = 1.0 + c * x;
            if (v > 0.0) break;
        }
        if (v <= 0.0) continue;
        v = v * v * v;
        const double u = s.uniform_d();
        if (u < 1.0 - 0.0331 * x * x * x * x || log(u) < 0.5 * x * x + d * (1.0 - v + log(v)))
            return d * v * boost * scale;
    }
    return d * boost * scale;
}

struct ReadoutArgs {
    shccd_params p;
    float* integrated;        // may be null
    const float* in_i;        // may be null: last integration step fused into the readout
    int64_t in_stride;
    double factor;
    const float* gains;       // may be null
    uint64_t counter;
    float* pixels;
    int64_t out_stride;
    int64_t int_stride;       // frame stride of the integrated buffer
    double inv_adu_gain, inv_emccd_gain;
};

// Integrated signal of pixel q of frame f without binning; reads and zeroes the integration
// buffer (integrated_i.i *= 0, ccd.py:217) and adds the fused last step in_i * factor (FP32, as
// Intensity.sum).
__device__ __forceinline__ float raw_pixel(const ReadoutArgs& a, int f, int q) {
    float acc = 0.0f;
    if (a.integrated) {
        float* ip = a.integrated + f * a.int_stride + q;
        acc = *ip;
        *ip = 0.0f;
    }
    if (a.in_i) acc += a.in_i[f * a.in_stride + q] * (float)a.factor;
    return acc;
}

// background + dark current added to every pixel (ccd.py:220-221), unless the input already
// carries them (SHCCD_OFFSET_APPLIED: the second stage of a CTE readout)
__device__ __forceinline__ float offset_of(const shccd_params& p) {
    return ((p.flags & (SHCCD_BACKGROUND | SHCCD_DARK)) && !(p.flags & SHCCD_OFFSET_APPLIED))
               ? (float)(p.background_level + p.darkcurrent_level) : 0.0f;
}

// apply_qe + background/dark (+ CIC in the FULL variants): the expected signal before the photon
// noise, from the integrated (binned) signal.
template <bool FULL>
__device__ __forceinline__ float signal_of(const ReadoutArgs& a, float acc, uint64_t px, uint32_t ctr) {
    const shccd_params& p = a.p;
    if (p.qe != 1.0) acc *= (float)p.qe;
    acc += offset_of(p);
    if (FULL && (p.flags & SHCCD_CIC)) {
        Stream s(p.cic_seed, px, ctr);
        if (p.cic_level > 0.0 && s.uniform_d() <= p.cic_level) acc += 1.0f;
    }
    return acc;
}

// apply_binning + signal_of for pixel q of frame f.
template <bool FULL>
__device__ __forceinline__ float load_pixel(const ReadoutArgs& a, int f, int q, uint64_t px, uint32_t ctr) {
    const shccd_params& p = a.p;
    float acc = 0.0f;
    if (!FULL || p.binning == 1) {
        acc = raw_pixel(a, f, q);
    } else {
        const int b = p.binning;
        const int r = q / p.out_cols, c = q - r * p.out_cols;
        for (int i = 0; i < b; ++i) {
            const int64_t row = (int64_t)(r * b + i) * p.in_cols + (int64_t)c * b;
            for (int j = 0; j < b; ++j) {
                float v = 0.0f;
                if (a.integrated) {
                    float* ip = a.integrated + f * a.int_stride + row + j;
                    v = *ip;
                    *ip = 0.0f;
                }
                if (a.in_i) v += a.in_i[f * a.in_stride + row + j] * (float)a.factor;
                acc += v;
            }
        }
        acc *= (float)(b * b);
    }
    return signal_of<FULL>(a, acc, px, ctr);
}

// Generator shared by photon and readout noise, keyed by both seeds.  Call 0 of a pixel gives the
// first Poisson try (words x, y) and the readout normal (words z, w); further tries use calls 1..
__device__ __forceinline__ uint32_t key0_of(const shccd_params& p) {
    return (uint32_t)p.photon_seed ^ (uint32_t)(p.readout_seed >> 32);
}
__device__ __forceinline__ uint32_t key1_of(const shccd_params& p) { return (uint32_t)p.readout_seed ^ 0x243F6A88u; }

// Everything after the photon noise: EM gain, readout noise (nrm ~ N(0,1)), pixel gains, ADU.
template <bool FULL>
__device__ __forceinline__ float finish_value(const ReadoutArgs& a, int q, uint64_t px, uint32_t ctr, double xd,
                                             bool integer_frame, float nrm) {
    const shccd_params& p = a.p;
    if (FULL && (p.flags & SHCCD_EXCESS)) {
        Stream s(p.excess_seed, px, ctr);
        xd = gamma_sample(s, p.excess_delta * xd, p.emccd_gain) / p.excess_delta;
        integer_frame = false;
    }
    if (p.flags & SHCCD_READOUT) {
        const float ron = nrm * (float)p.readout_level;
        xd += integer_frame ? trunc((double)ron) : (double)ron;
    }
    if (FULL && a.gains) xd *= (double)a.gains[q];
    if (p.flags & SHCCD_PHOTON) {
        xd = rint(xd * p.adu_gain) + p.adu_bias;                   // round half to even (numpy round)
        if (xd < 0.0) xd = 0.0;
        if (!(p.flags & SHCCD_KEEP_BIAS)) xd -= p.adu_bias;
        xd *= a.inv_adu_gain;           // /ADU_gain: exact for power-of-two gains, else within 1 ulp
        if (FULL && (p.flags & SHCCD_EXCESS)) xd *= a.inv_emccd_gain;
        if ((p.flags & SHCCD_DARK) && !(p.flags & SHCCD_KEEP_DARK)) xd -= p.darkcurrent_level;
        if ((p.flags & SHCCD_BG_REMOVE_AVG) && !(p.flags & SHCCD_KEEP_DARK)) xd -= p.background_level;
    }
    return (float)xd;
}

// Readout without photon noise: groups of 4 pixels per lane (float4 traffic when the frame
// allows it); one generator call per group gives 4 readout normals (two Box-Muller pairs).
template <bool FULL>
__global__ void __launch_bounds__(256) ccd_readout_kernel(ReadoutArgs a, int vec) {
    const int f = blockIdx.y;
    const int n_out = a.p.out_rows * a.p.out_cols;
    const int n_grp = (n_out + 3) >> 2;
    const uint32_t ctr = (uint32_t)a.counter;
    const uint32_t k0 = key0_of(a.p), k1 = key1_of(a.p);
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < n_grp; g += gridDim.x * blockDim.x) {
        const int q0 = g << 2;
        float x[4];
        if (vec) {
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (a.integrated) {
                float4* ip = reinterpret_cast<float4*>(a.integrated + f * a.int_stride + q0);
                v = *ip;
                *ip = make_float4(0.f, 0.f, 0.f, 0.f);             // integrated_i.i *= 0 (ccd.py:217)
            }
            if (a.in_i) {
                const float4 w = *reinterpret_cast<const float4*>(a.in_i + f * a.in_stride + q0);
                const float fc = (float)a.factor;
                v.x += w.x * fc; v.y += w.y * fc; v.z += w.z * fc; v.w += w.w * fc;
            }
            x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
            const float qe = (float)a.p.qe;
            const float add = offset_of(a.p);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (a.p.qe != 1.0) x[j] *= qe;
                x[j] += add;
                if (FULL && (a.p.flags & SHCCD_CIC)) {
                    Stream s(a.p.cic_seed, (uint64_t)f * (uint64_t)n_out + (uint64_t)(q0 + j), ctr);
                    if (a.p.cic_level > 0.0 && s.uniform_d() <= a.p.cic_level) x[j] += 1.0f;
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                x[j] = q0 + j < n_out ? load_pixel<FULL>(a, f, q0 + j, (uint64_t)f * (uint64_t)n_out + (uint64_t)(q0 + j), ctr)
                                      : 0.0f;
        }
        float nrm[4] = {0.f, 0.f, 0.f, 0.f};
        if (a.p.flags & SHCCD_READOUT) {
            const uint64_t gid = (uint64_t)f * (uint64_t)n_grp + (uint64_t)g;
            const U4 w = philox(U4{(uint32_t)gid, (uint32_t)(gid >> 32), ctr, 0x80000000u}, k0, k1);
            const float r0 = __builtin_sqrtf(-2.0f * fast_log(u01(w.x))), t0 = u01(w.y);
            const float r1 = __builtin_sqrtf(-2.0f * fast_log(u01(w.z))), t1 = u01(w.w);
            nrm[0] = r0 * __builtin_amdgcn_cosf(t0);
            nrm[1] = r0 * __builtin_amdgcn_sinf(t0);
            nrm[2] = r1 * __builtin_amdgcn_cosf(t1);
            nrm[3] = r1 * __builtin_amdgcn_sinf(t1);
        }
        float y[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            y[j] = finish_value<FULL>(a, q0 + j < n_out ? q0 + j : q0, (uint64_t)f * (uint64_t)n_out + (uint64_t)(q0 + j), ctr,
                                (double)x[j], false, nrm[j]);
        if (vec) {
            *reinterpret_cast<float4*>(a.pixels + f * a.out_stride + q0) = make_float4(y[0], y[1], y[2], y[3]);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (q0 + j < n_out) a.pixels[f * a.out_stride + q0 + j] = y[j];
        }
    }
}

// Samplers of the Poisson photon count, shared by the photon kernels.  PTRS constants of a pixel
// (lam >= 10), in Hormann's notation.
struct PtrsConst {
    float b, a, log_invalpha, vr, log_lam, inv_lam;
};

__device__ __forceinline__ PtrsConst ptrs_const(float lam) {
    PtrsConst c;
    const float slam = __builtin_amdgcn_sqrtf(lam);
    c.b = 0.931f + 2.53f * slam;
    c.a = -0.059f + 0.02483f * c.b;
    c.log_invalpha = fast_log(1.1239f + 1.1328f * fast_rcp(c.b - 3.4f));
    c.vr = 0.9277f - 3.6224f * fast_rcp(c.b - 2.0f);
    c.log_lam = fast_log(lam);
    c.inv_lam = fast_rcp(lam);
    return c;
}

// One PTRS try with the uniforms u1, u2 in (0, 1]: the candidate count k, true when accepted.
__device__ __forceinline__ bool ptrs_try(const PtrsConst& c, float lam, float u1, float u2, double& k) {
    const float U = u1 - 0.5f, V = u2;
    const float us = 0.5f - fabsf(U);
    k = floor((double)((2.0f * c.a * fast_rcp(us) + c.b) * U) + (double)lam + 0.43);
    bool done = us >= 0.07f && V <= c.vr;
    if (!done && !(k < 0.0 || (us < 0.013f && V > us))) {
        const float lhs = fast_log(V) + c.log_invalpha - fast_log(c.a * fast_rcp(us * us) + c.b);
        const float rhs = k < 32.0 ? -lam + (float)k * c.log_lam - (float)kLogFact[(int)k]
                                   : log_pmf_large(k, lam, c.inv_lam);
        done = lhs <= rhs;
    }
    return done;
}

// Inversion of the CDF for 0 < lam < 10 with one 64-bit uniform made of the words w0, w1 (the
// search ends within 100 steps unless U is within 1e-40 of 1).
__device__ __forceinline__ double poisson_inversion(float lam, uint32_t w0, uint32_t w1) {
    const float enlam = __builtin_amdgcn_exp2f(-1.44269504089f * lam);
    const double U = ((double)w0 * 4294967296.0 + (double)w1 + 1.0) * (1.0 / 18446744073709551616.0);
    double pk = (double)enlam, F = pk, k = 0.0;
    for (int s = 0; s < 100 && U > F; ++s) {
        k += 1.0;
        pk *= (double)(lam * fast_rcp((float)k));
        F += pk;
    }
    return k;
}

// Readout with photon noise.  Every pixel draws from the generator keyed by (pixel, readout counter,
// try): try 0 gives the inversion uniform or the first PTRS try (words x, y) and the readout normal
// (words z, w), further PTRS tries use calls 1, 2, ...  So a pixel's value does not depend on which
// lane samples it or in which order, and the two kernels below give the same values bit for bit.
//
// Each wave owns contiguous tiles of kPhotonTile pixels (tile t of the frame, waves striding over
// the tiles), parks each pixel's value in its LDS slice when the pixel is done, and stores the tile
// with coalesced whole-line writes once every lane is through (the lanes finish their pixels at
// different times, so storing each value when it is done would write partial cache lines).
constexpr int kPhotonPerLane = 16;
constexpr int kPhotonTile = 64 * kPhotonPerLane;

__device__ __forceinline__ void ccd_wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// rank of this lane among the lanes set in the wave mask m
__device__ __forceinline__ int lane_rank(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Full variant (binning, CIC, EM gain, pixel gains): lane l walks pixels l, l + 64, ... of the
// tile, one generator call per loop iteration, and a lane that accepts moves to its next pixel while
// the others keep trying.
__global__ void __launch_bounds__(256) ccd_photon_full_kernel(ReadoutArgs a) {
    __shared__ float stage[4][kPhotonTile];
    const shccd_params& p = a.p;
    const int f = blockIdx.y;
    const int n_out = p.out_rows * p.out_cols;
    const uint32_t ctr = (uint32_t)a.counter;
    const uint32_t k0 = key0_of(p), k1 = key1_of(p);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float* st = stage[wave];
    float* out = a.pixels + f * a.out_stride;
    const int n_tiles = (n_out + kPhotonTile - 1) / kPhotonTile;
    for (int tile = blockIdx.x * 4 + wave; tile < n_tiles; tile += gridDim.x * 4) {
        const int t0 = tile * kPhotonTile;
        const int tn = min(kPhotonTile, n_out - t0);
        int i = lane;                          // pixel t0 + i of the tile
        bool fresh = true;
        uint64_t px = 0;
        uint32_t wz = 0, ww = 0;
        float lam = 0.0f;
        PtrsConst pc{};
        double k = 0.0;
        uint32_t tries = 0;
        while (i < tn) {
            const int q = t0 + i;
            if (fresh) {
                px = (uint64_t)f * (uint64_t)n_out + (uint64_t)q;
                lam = load_pixel<true>(a, f, q, px, ctr);
                tries = 0;
                k = 0.0;
                pc = ptrs_const(lam);
                fresh = false;
            }
            const U4 w = philox(U4{(uint32_t)px, (uint32_t)(px >> 32), ctr, tries}, k0, k1);
            if (tries == 0) {
                wz = w.z;
                ww = w.w;
            }
            ++tries;
            bool done;
            if (!(lam > 0.0f)) {
                k = 0.0;
                done = true;
            } else if (lam < 10.0f) {
                k = poisson_inversion(lam, w.x, w.y);
                done = true;
            } else {
                done = ptrs_try(pc, lam, u01(w.x), u01(w.y), k);
            }
            if (!done && tries >= 512u) {                              // never reached in practice
                k = fmax(k, 0.0);
                done = true;
            }
            if (done) {
                st[i] = finish_value<true>(a, q, px, ctr, k, true, normal_from(wz, ww));
                i += 64;
                fresh = true;
            }
        }
        ccd_wave_lds_sync();                   // every lane of the wave has parked its pixels
        for (int j = lane; j < tn; j += 64) out[t0 + j] = st[j];
        ccd_wave_lds_sync();                   // the slice is read before the next tile reuses it
    }
}

// Lean variant (the common case: no binning, CIC, EM gain or pixel gains).  A wave that mixes
// pixels without signal (k = 0), pixels sampled by inversion (0 < lam < 10) and PTRS pixels pays for
// all three code paths on every try, so the tile's pixels are first listed by class (wave ballots
// into an LDS list) and each class then runs with every lane on it.  PTRS lanes take the next list
// entry as soon as they accept, so a lane whose pixel is rejected does not hold the others back.
// `vec`: 16-byte aligned frames, the tile is stored with float4 writes.
__global__ void __launch_bounds__(256) ccd_photon_lean_kernel(ReadoutArgs a, int vec) {
    __shared__ __align__(16) float val[4][kPhotonTile];   // signal of each tile pixel, then its value
    __shared__ uint16_t order[4][kPhotonTile];     // tile pixels by class: zero | inversion | PTRS
    const shccd_params& p = a.p;
    const int f = blockIdx.y;
    const int n_out = p.out_rows * p.out_cols;
    const uint32_t ctr = (uint32_t)a.counter;
    const uint32_t k0 = key0_of(p), k1 = key1_of(p);
    const bool readout = (p.flags & SHCCD_READOUT) != 0;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float* vw = val[wave];
    uint16_t* ow = order[wave];
    float* out = a.pixels + f * a.out_stride;
    const uint64_t fbase = (uint64_t)f * (uint64_t)n_out;
    const int n_tiles = (n_out + kPhotonTile - 1) / kPhotonTile;
    for (int tile = blockIdx.x * 4 + wave; tile < n_tiles; tile += gridDim.x * 4) {
        const int t0 = tile * kPhotonTile;
        const int tn = min(kPhotonTile, n_out - t0);
        // signals of the lane's pixels i = lane + 64 j, every load in flight before the first use
        // (slots past the end of a ragged last tile read its last pixel and are marked invalid)
        const bool whole = tn == kPhotonTile;
        float raw[kPhotonPerLane];
#pragma unroll
        for (int j = 0; j < kPhotonPerLane; ++j) raw[j] = 0.0f;
        if (a.integrated) {
            float* ip = a.integrated + f * a.int_stride + t0;
#pragma unroll
            for (int j = 0; j < kPhotonPerLane; ++j) raw[j] = ip[min(lane + 64 * j, tn - 1)];
#pragma unroll
            for (int j = 0; j < kPhotonPerLane; ++j)
                if (whole || lane + 64 * j < tn) ip[lane + 64 * j] = 0.0f;       // integrated_i.i *= 0 (ccd.py:217)
        }
        if (a.in_i) {
            const float* src = a.in_i + f * a.in_stride + t0;
            const float fc = (float)a.factor;
            float w[kPhotonPerLane];
#pragma unroll
            for (int j = 0; j < kPhotonPerLane; ++j) w[j] = src[min(lane + 64 * j, tn - 1)];
#pragma unroll
            for (int j = 0; j < kPhotonPerLane; ++j) raw[j] += w[j] * fc;
        }
        // class of each pixel (2 bits per slot j: 0 no signal, 1 inversion, 2 PTRS, 3 past the end)
        uint32_t cls = 0;
        int c0 = 0, c1 = 0;
#pragma unroll
        for (int j = 0; j < kPhotonPerLane; ++j) {
            const int i = lane + 64 * j;
            const float lam = signal_of<false>(a, raw[j], fbase + (uint64_t)(t0 + i), ctr);
            vw[i] = lam;
            const uint32_t c = !(whole || i < tn) ? 3u : !(lam > 0.0f) ? 0u : lam < 10.0f ? 1u : 2u;
            cls |= c << (2 * j);
            c0 += c == 0u;
            c1 += c == 1u;
        }
        // list positions: the lane's pixels of each class follow those of the lower lanes (inclusive
        // scan over the wave of the packed counts c0 + 2^16 c1; totals up to 1024 fit in 16 bits)
        const uint32_t own = (uint32_t)c0 | ((uint32_t)c1 << 16);
        uint32_t scan = own;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(scan, o, 64);
            if (lane >= o) scan += v;
        }
        const uint32_t tot = __shfl(scan, 63, 64);
        const int n0 = (int)(tot & 0xffffu), n1 = (int)(tot >> 16);
        const uint32_t ex = scan - own;
        const int valid_before = whole ? 16 * lane : lane * (tn / 64) + min(lane, tn % 64);
        int e0 = (int)(ex & 0xffffu), e1 = n0 + (int)(ex >> 16);
        int e2 = n0 + n1 + valid_before - (int)(ex & 0xffffu) - (int)(ex >> 16);
#pragma unroll
        for (int j = 0; j < kPhotonPerLane; ++j) {
            const uint32_t c = (cls >> (2 * j)) & 3u;
            const int pos = c == 0u ? e0 : c == 1u ? e1 : e2;
            if (c != 3u) ow[pos] = (uint16_t)(lane + 64 * j);
            e0 += c == 0u;
            e1 += c == 1u;
            e2 += c == 2u;
        }
        const int n_all = whole ? kPhotonTile : tn;
        ccd_wave_lds_sync();
        // no signal: k = 0 (the generator only for the readout normal)
        for (int e = lane; e < n0; e += 64) {
            const int i = ow[e], q = t0 + i;
            const uint64_t px = fbase + (uint64_t)q;
            float nrm = 0.0f;
            if (readout) {
                const U4 w = philox(U4{(uint32_t)px, (uint32_t)(px >> 32), ctr, 0u}, k0, k1);
                nrm = normal_from(w.z, w.w);
            }
            vw[i] = finish_value<false>(a, q, px, ctr, 0.0, true, nrm);
        }
        // 0 < lam < 10: inversion of the CDF, one generator call
        for (int e = n0 + lane; e < n0 + n1; e += 64) {
            const int i = ow[e], q = t0 + i;
            const uint64_t px = fbase + (uint64_t)q;
            const U4 w = philox(U4{(uint32_t)px, (uint32_t)(px >> 32), ctr, 0u}, k0, k1);
            const double k = poisson_inversion(vw[i], w.x, w.y);
            vw[i] = finish_value<false>(a, q, px, ctr, k, true, normal_from(w.z, w.w));
        }
        // lam >= 10: PTRS, one try per iteration; a lane that accepts takes the next entry
        {
            int e = n0 + n1 + lane;                    // this lane's entry (>= n_all: none left)
            int next = n0 + n1 + 64;                   // first entry nobody has taken
            bool fresh = true;
            int i = 0;
            uint64_t px = 0;
            float lam = 0.0f;
            PtrsConst pc{};
            uint32_t tries = 0, wz = 0, ww = 0;
            while (__any(e < n_all)) {
                bool done = false;
                if (e < n_all) {
                    if (fresh) {
                        i = ow[e];
                        px = fbase + (uint64_t)(t0 + i);
                        lam = vw[i];
                        pc = ptrs_const(lam);
                        tries = 0;
                        fresh = false;
                    }
                    const U4 w = philox(U4{(uint32_t)px, (uint32_t)(px >> 32), ctr, tries}, k0, k1);
                    if (tries == 0) {
                        wz = w.z;
                        ww = w.w;
                    }
                    ++tries;
                    double k;
                    done = ptrs_try(pc, lam, u01(w.x), u01(w.y), k);
                    if (!done && tries >= 512u) {                  // never reached in practice
                        k = fmax(k, 0.0);
                        done = true;
                    }
                    if (done) {
                        vw[i] = finish_value<false>(a, t0 + i, px, ctr, k, true, normal_from(wz, ww));
                        fresh = true;
                    }
                }
                const uint64_t dm = __ballot(done);
                if (done) e = next + lane_rank(dm);
                next += __popcll(dm);
            }
        }
        ccd_wave_lds_sync();                   // every pixel of the tile has its value
        if (vec && tn == kPhotonTile) {
#pragma unroll
            for (int j = 0; j < kPhotonTile / 256; ++j) {
                const int o = 4 * (lane + 64 * j);
                *reinterpret_cast<float4*>(out + t0 + o) = *reinterpret_cast<const float4*>(vw + o);
            }
        } else {
            for (int j = lane; j < tn; j += 64) out[t0 + j] = vw[j];
        }
        ccd_wave_lds_sync();                   // the slice is read before the next tile reuses it
    }
}

__global__ void __launch_bounds__(256) ccd_integrate_kernel(const float* __restrict__ in_i, int64_t in_stride,
                                                            double factor, float* __restrict__ integrated,
                                                            int64_t int_stride, int64_t n_px) {
    const int f = blockIdx.y;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_px; q += (int64_t)gridDim.x * blockDim.x)
        integrated[f * int_stride + q] += in_i[f * in_stride + q] * (float)factor;
}

// blocks of the photon kernel: one tile of kPhotonTile pixels per wave, 4 waves per block
int photon_blocks(int64_t n) {
    const int64_t tiles = (n + kPhotonTile - 1) / kPhotonTile;
    return (int)std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4, 4096));
}

// blocks for n pixels with at least `per_lane` pixels per lane (capped at 4096 blocks)
int grid_for(int64_t n, int per_lane) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + 256LL * per_lane - 1) / (256LL * per_lane), 4096));
}

}  // namespace

extern "C" {

int shccd_philox(int n, const uint32_t* counters_keys, uint32_t* out, void* stream) {
    if (n < 0 || (n > 0 && (!counters_keys || !out)))
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_philox: bad argument");
    if (n == 0) return SHWFS_OK;
    hipLaunchKernelGGL(philox_kat_kernel, dim3((n + 255) / 256), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       counters_keys, out, n);
    CCD_HIP_CHECK(hipGetLastError());
    return SHWFS_OK;
}

int shccd_integrate(int n_frames, const float* in_i, int64_t in_stride, double factor, float* integrated,
                    int64_t int_stride, int64_t n_px, void* stream) {
    if (!in_i || !integrated || n_frames <= 0 || n_frames > 65535 || n_px <= 0)
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_integrate: bad argument");
    if (n_frames > 1 && (in_stride < n_px || int_stride < n_px))
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_integrate: frame stride too small");
    hipLaunchKernelGGL(ccd_integrate_kernel, dim3(grid_for(n_px, 1), n_frames), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), in_i, in_stride, factor, integrated, int_stride, n_px);
    CCD_HIP_CHECK(hipGetLastError());
    return SHWFS_OK;
}

int shccd_readout(const shccd_params* p, int n_frames, float* integrated, int64_t int_stride, const float* in_i,
                  int64_t in_stride, double factor, const float* pixel_gains, uint64_t counter, float* pixels,
                  int64_t out_stride, void* stream) {
    if (!p || !pixels || n_frames <= 0 || n_frames > 65535 || (!integrated && !in_i))
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_readout: bad argument (null buffers or n_frames outside [1, 65535])");
    if (p->binning < 1 || p->out_rows <= 0 || p->out_cols <= 0 || p->in_rows != p->out_rows * p->binning ||
        p->in_cols != p->out_cols * p->binning)
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_readout: frame sizes must satisfy in = out * binning");
    if ((p->flags & SHCCD_PHOTON) && !(p->adu_gain > 0.0))
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_readout: ADU gain must be > 0");
    if ((p->flags & SHCCD_EXCESS) && !(p->excess_delta > 0.0 && p->emccd_gain > 0.0))
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_readout: excess noise needs delta > 0 and EM gain > 0");
    const int64_t n_in = (int64_t)p->in_rows * p->in_cols, n_out = (int64_t)p->out_rows * p->out_cols;
    if (n_frames > 1 && ((integrated && int_stride < n_in) || (in_i && in_stride < n_in) || out_stride < n_out))
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_readout: frame stride too small");
    ReadoutArgs a{};
    a.p = *p;
    a.integrated = integrated;
    a.int_stride = int_stride;
    a.in_i = in_i;
    a.in_stride = in_stride;
    a.factor = factor;
    a.gains = pixel_gains;
    a.counter = counter;
    a.pixels = pixels;
    a.out_stride = out_stride;
    a.inv_adu_gain = (p->flags & SHCCD_PHOTON) ? 1.0 / p->adu_gain : 1.0;
    a.inv_emccd_gain = (p->flags & SHCCD_EXCESS) ? 1.0 / p->emccd_gain : 1.0;
    if (n_out > (int64_t)INT32_MAX) return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_readout: frame too large");
    // FULL variants carry binning, CIC, EM gain and pixel gains; the lean ones the common rest
    const bool full = p->binning != 1 || (p->flags & (SHCCD_CIC | SHCCD_EXCESS)) || pixel_gains != nullptr;
    if (p->flags & SHCCD_PHOTON) {
        if (full) {
            hipLaunchKernelGGL(ccd_photon_full_kernel, dim3(photon_blocks(n_out), n_frames), dim3(256), 0,
                               reinterpret_cast<hipStream_t>(stream), a);
        } else {
            auto al = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15u) == 0; };
            const int vec = al(pixels) && (n_frames == 1 || out_stride % 4 == 0);
            hipLaunchKernelGGL(ccd_photon_lean_kernel, dim3(photon_blocks(n_out), n_frames), dim3(256), 0,
                               reinterpret_cast<hipStream_t>(stream), a, vec);
        }
    } else {
        auto al = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15u) == 0; };
        const int vec = p->binning == 1 && n_out % 4 == 0 && al(pixels) && (!integrated || al(integrated)) &&
                        (!in_i || al(in_i)) &&
                        (n_frames == 1 || (out_stride % 4 == 0 && int_stride % 4 == 0 && in_stride % 4 == 0));
        if (full)
            hipLaunchKernelGGL(ccd_readout_kernel<true>, dim3(grid_for((n_out + 3) / 4, 4), n_frames), dim3(256), 0,
                               reinterpret_cast<hipStream_t>(stream), a, vec);
        else
            hipLaunchKernelGGL(ccd_readout_kernel<false>, dim3(grid_for((n_out + 3) / 4, 4), n_frames), dim3(256), 0,
                               reinterpret_cast<hipStream_t>(stream), a, vec);
    }
    CCD_HIP_CHECK(hipGetLastError());
    return SHWFS_OK;
}


// CTE readout (ccd.py:220-224): frame = cte_rows . (binned QE signal + background + dark) . cte_cols,
// then CIC, photon, EM, readout noise, gains and ADU conversion as shccd_readout.  Three stages on
// the stream: the readout kernel without noise writes the offset signal into work[0]; two FP32
// matrix products (shgemm_f32) form the transferred frame in work[1] and back in work[0]; the
// readout kernel with the noise flags and SHCCD_OFFSET_APPLIED (background and dark are not added
// again; their subtraction after the ADU conversion is kept) reads it into `pixels`.
int shccd_readout_cte(const shccd_params* p, int n_frames, float* integrated, int64_t int_stride, const float* in_i,
                      int64_t in_stride, double factor, const float* cte_rows, const float* cte_cols, float* work,
                      const float* pixel_gains, uint64_t counter, float* pixels, int64_t out_stride, void* stream) {
    if (!p || !cte_rows || !cte_cols || !work)
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_readout_cte: null CTE matrices or work buffer");
    if (n_frames <= 0 || n_frames > 65535)
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_readout_cte: n_frames outside [1, 65535]");
    const int R = p->out_rows, C = p->out_cols;
    const int64_t n_out = (int64_t)R * C;
    if (R <= 0 || C <= 0) return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_readout_cte: empty frame");
    float* w0 = work;                               // [n_frames][R][C]
    float* w1 = work + (int64_t)n_frames * n_out;   // [n_frames][R][C]
    // stage 1: binning, QE, background + dark; no noise, no gains
    shccd_params s1 = *p;
    s1.flags = p->flags & (SHCCD_BACKGROUND | SHCCD_DARK);
    int rc = shccd_readout(&s1, n_frames, integrated, int_stride, in_i, in_stride, factor, nullptr, counter, w0,
                           n_out, stream);
    if (rc != SHWFS_OK) return rc;
    // stage 2: w1[f] = cte_rows . w0[f] (R x R . R x C), then w0 = w1 . cte_cols over all frames at once
    for (int f = 0; f < n_frames; ++f) {
        rc = shgemm_f32(R, C, R, 1.0f, cte_rows, R, w0 + (int64_t)f * n_out, C, 0, 0.0f, nullptr,
                        w1 + (int64_t)f * n_out, C, stream);
        if (rc != SHWFS_OK) return rc;
    }
    if ((int64_t)n_frames * R > INT32_MAX)
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_readout_cte: too many rows for one product");
    rc = shgemm_f32(n_frames * R, C, C, 1.0f, w1, C, cte_cols, C, 0, 0.0f, nullptr, w0, C, stream);
    if (rc != SHWFS_OK) return rc;
    // stage 3: the noise chain on the transferred frame
    shccd_params s3 = *p;
    s3.in_rows = R;
    s3.in_cols = C;
    s3.binning = 1;
    s3.qe = 1.0;
    s3.flags = p->flags | SHCCD_OFFSET_APPLIED;
    return shccd_readout(&s3, n_frames, nullptr, 0, w0, n_out, 1.0, pixel_gains, counter, pixels, out_stride, stream);
}

}  // extern "C"
