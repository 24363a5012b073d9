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
// Poisson sampling: multiplication of uniforms (Knuth) for lambda < 10, Hormann's transformed
// rejection with squeeze (PTRS, Insurance: Mathematics and Economics 12 (1993) 39-45) above,
// with the acceptance test in FP64.  Gamma: Marsaglia & Tsang (ACM TOMS 26 (2000) 363-372).

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

__device__ __forceinline__ U4 philox(U4 c, uint32_t k0, uint32_t k1) {
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
        const uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Stream of uniforms of one pixel for one noise source: draw d -> 4 words per Philox call.
struct Stream {
    uint32_t k0, k1;          // seed
    uint32_t px_lo, px_hi;    // pixel index
    uint32_t ctr;             // readout counter
    uint32_t draw = 0;
    U4 buf{};
    int left = 0;

    __device__ Stream(uint64_t seed, uint64_t px, uint32_t counter)
        : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), px_lo((uint32_t)px), px_hi((uint32_t)(px >> 32)),
          ctr(counter) {}

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

// ln(k!) = lnGamma(k + 1): table below 32, Stirling's series (to x^-7) above (error < 1e-16).
__device__ __forceinline__ double log_factorial(double k) {
    if (k < 32.0) return kLogFact[(int)k];
    const double x = k + 1.0, r = 1.0 / x, r2 = r * r;
    const double series = r * (1.0 / 12 - r2 * (1.0 / 360 - r2 * (1.0 / 1260 - r2 * (1.0 / 1680))));
    return (x - 0.5) * log(x) - x + 0.91893853320467274178 + series;
}

__device__ double poisson(Stream& s, double lam) {
    if (!(lam > 0.0)) return 0.0;
    if (lam < 10.0) {                                     // product of uniforms vs exp(-lam)
        const double enlam = exp(-lam);
        double prod = 1.0;
        int k = 0;
        for (int it = 0; it < 1024; ++it) {               // P(k > 1023 | lam < 10) = 0 in double
            prod *= s.uniform_d();
            if (prod <= enlam) break;
            ++k;
        }
        return (double)k;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    double k = floor(lam);
    for (int it = 0; it < 256; ++it) {                    // acceptance > 0.85 per try
        const double U = s.uniform_d() - 0.5;
        const double V = s.uniform_d();
        const double us = 0.5 - fabs(U);
        k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - log_factorial(k)) return k;
    }
    return fmax(k, 0.0);
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
};

__global__ void __launch_bounds__(256) ccd_readout_kernel(ReadoutArgs a) {
    const shccd_params& p = a.p;
    const int f = blockIdx.y;
    const int64_t n_out = (int64_t)p.out_rows * p.out_cols;
    const int b = p.binning;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_out; q += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(q / p.out_cols), c = (int)(q - (int64_t)r * p.out_cols);
        float acc = 0.0f;
        for (int i = 0; i < b; ++i) {
            const int64_t row = (int64_t)(r * b + i) * p.in_cols + (int64_t)c * b;
            for (int j = 0; j < b; ++j) {
                float v = 0.0f;
                if (a.integrated) {
                    float* ip = a.integrated + f * a.int_stride + row + j;
                    v = *ip;
                    *ip = 0.0f;                                   // integrated_i.i *= 0 (ccd.py:217)
                }
                if (a.in_i) v += a.in_i[f * a.in_stride + row + j] * (float)a.factor;   // FP32 as Intensity.sum
                acc += v;
            }
        }
        float x = acc * (float)(b * b);
        if (p.qe != 1.0) x *= (float)p.qe;
        const uint64_t px = (uint64_t)f * (uint64_t)n_out + (uint64_t)q;
        const uint32_t ctr = (uint32_t)a.counter;
        if (p.flags & (SHCCD_BACKGROUND | SHCCD_DARK)) x += (float)(p.background_level + p.darkcurrent_level);
        double xd = x;
        if (p.flags & SHCCD_CIC) {
            Stream s(p.cic_seed, px, ctr);
            xd += (s.uniform_d() <= p.cic_level && p.cic_level > 0.0) ? 1.0 : 0.0;
        }
        bool integer_frame = false;
        if (p.flags & SHCCD_PHOTON) {
            Stream s(p.photon_seed, px, ctr);
            xd = poisson(s, xd);
            integer_frame = true;
        }
        if (p.flags & SHCCD_EXCESS) {
            Stream s(p.excess_seed, px, ctr);
            xd = gamma_sample(s, p.excess_delta * xd, p.emccd_gain) / p.excess_delta;
            integer_frame = false;
        }
        if (p.flags & SHCCD_READOUT) {
            Stream s(p.readout_seed, px, ctr);
            const double ron = (double)s.normal() * p.readout_level;
            xd += integer_frame ? trunc(ron) : (double)(float)ron;
        }
        if (a.gains) xd *= (double)a.gains[q];
        if (p.flags & SHCCD_PHOTON) {
            xd = rint(xd * p.adu_gain) + p.adu_bias;          // round half to even (numpy round)
            if (xd < 0.0) xd = 0.0;
            if (!(p.flags & SHCCD_KEEP_BIAS)) xd -= p.adu_bias;
            xd /= p.adu_gain;
            if (p.flags & SHCCD_EXCESS) xd /= p.emccd_gain;
            if ((p.flags & SHCCD_DARK) && !(p.flags & SHCCD_KEEP_DARK)) xd -= p.darkcurrent_level;
            if ((p.flags & SHCCD_BG_REMOVE_AVG) && !(p.flags & SHCCD_KEEP_DARK)) xd -= p.background_level;
        }
        a.pixels[f * a.out_stride + q] = (float)xd;
    }
}

__global__ void __launch_bounds__(256) ccd_integrate_kernel(const float* __restrict__ in_i, int64_t in_stride,
                                                            double factor, float* __restrict__ integrated,
                                                            int64_t int_stride, int64_t n_px) {
    const int f = blockIdx.y;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_px; q += (int64_t)gridDim.x * blockDim.x)
        integrated[f * int_stride + q] += in_i[f * in_stride + q] * (float)factor;
}

int grid_for(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 4096); }

}  // namespace

extern "C" {

int shccd_integrate(int n_frames, const float* in_i, int64_t in_stride, double factor, float* integrated,
                    int64_t int_stride, int64_t n_px, void* stream) {
    if (!in_i || !integrated || n_frames <= 0 || n_px <= 0)
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_integrate: bad argument");
    if (n_frames > 1 && (in_stride < n_px || int_stride < n_px))
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_integrate: frame stride too small");
    hipLaunchKernelGGL(ccd_integrate_kernel, dim3(grid_for(n_px), n_frames), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), in_i, in_stride, factor, integrated, int_stride, n_px);
    CCD_HIP_CHECK(hipGetLastError());
    return SHWFS_OK;
}

int shccd_readout(const shccd_params* p, int n_frames, float* integrated, int64_t int_stride, const float* in_i,
                  int64_t in_stride, double factor, const float* pixel_gains, uint64_t counter, float* pixels,
                  int64_t out_stride, void* stream) {
    if (!p || !pixels || n_frames <= 0 || (!integrated && !in_i))
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "shccd_readout: bad argument");
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
    hipLaunchKernelGGL(ccd_readout_kernel, dim3(grid_for(n_out), n_frames), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), a);
    CCD_HIP_CHECK(hipGetLastError());
    return SHWFS_OK;
}

}  // extern "C"
