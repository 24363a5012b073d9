// fft_reg.h -- register-resident radix-2 FFTs for gfx950, fully unrolled at compile time.
//
// One work-item owns one whole transform: the butterflies never leave its VGPRs, so no
// cross-lane traffic is needed.  Twiddles are compile-time constants (constexpr Taylor
// series in FP64, rounded once to FP32) that fold into instruction immediates; the trivial
// twiddles 1 and -+i are special-cased so they cost no multiplies.
//
// Convention (numpy.fft):  forward X[k] = sum_n x[n] exp(-2 pi i n k / L),
//                          inverse X[k] = sum_n x[n] exp(+2 pi i n k / L)  (unnormalised).
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <utility>

namespace shwfs {

constexpr double kPi = 3.14159265358979323846264338327950288;

// sin/cos by Taylor series on [-pi, pi]; 30 terms give full FP64 accuracy there.
constexpr double cx_sin(double x) {
    double term = x, sum = x;
    for (int n = 1; n < 30; ++n) { term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0)); sum += term; }
    return sum;
}
constexpr double cx_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 30; ++n) { term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n)); sum += term; }
    return sum;
}
// angle of exp(-+2 pi i k / L) reduced to [-pi, pi]
constexpr double tw_angle(int k, int L, bool inv) {
    int kk = k % L;
    double a = 2.0 * kPi * (double)kk / (double)L;   // in [0, 2 pi)
    if (a > kPi) a -= 2.0 * kPi;
    return inv ? a : -a;
}
constexpr float tw_re(int k, int L, bool inv) { return (float)cx_cos(tw_angle(k, L, inv)); }
constexpr float tw_im(int k, int L, bool inv) { return (float)cx_sin(tw_angle(k, L, inv)); }

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }
constexpr int bitrev(int i, int bits) {
    int r = 0;
    for (int b = 0; b < bits; ++b) { r = (r << 1) | (i & 1); i >>= 1; }
    return r;
}

template <int Begin, int End, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (Begin < End) {
        f(std::integral_constant<int, Begin>{});
        static_for<Begin + 1, End>(f);
    }
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// x * exp(-+2 pi i k / L) with trivial cases folded at compile time.
template <int K, int L, bool INV>
__device__ __forceinline__ float2 twiddle(float2 x) {
    constexpr int k = ((K % L) + L) % L;
    if constexpr (k == 0) {
        return x;
    } else if constexpr (4 * k == L) {            // -i (forward) / +i (inverse)
        return INV ? make_float2(-x.y, x.x) : make_float2(x.y, -x.x);
    } else if constexpr (2 * k == L) {
        return make_float2(-x.x, -x.y);
    } else if constexpr (4 * k == 3 * L) {        // +i (forward) / -i (inverse)
        return INV ? make_float2(x.y, -x.x) : make_float2(-x.y, x.x);
    } else {
        constexpr float wr = tw_re(k, L, INV);
        constexpr float wi = tw_im(k, L, INV);
        return make_float2(x.x * wr - x.y * wi, x.x * wi + x.y * wr);
    }
}

// In-place radix-2 decimation-in-time FFT of length L (power of two), natural order in/out.
template <int L, bool INV>
__device__ __forceinline__ void fft_radix2(float2 (&x)[L]) {
    constexpr int bits = ilog2(L);
    static_for<0, L>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        constexpr int r = bitrev(i, bits);
        if constexpr (r > i) { float2 t = x[i]; x[i] = x[r]; x[r] = t; }
    });
    static_for<1, bits + 1>([&](auto sc) {
        constexpr int m = 1 << decltype(sc)::value;
        constexpr int h = m / 2;
        static_for<0, L / m>([&](auto bc) {
            constexpr int base = decltype(bc)::value * m;
            static_for<0, h>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                float2 t = twiddle<j * (L / m), L, INV>(x[base + j + h]);
                float2 u = x[base + j];
                x[base + j] = cadd(u, t);
                x[base + j + h] = csub(u, t);
            });
        });
    });
}

// Length-M transform of an input whose samples n >= NP are zero, NP <= M/2.
// First decimation-in-frequency stage done by hand: the upper half of the input is zero, so
// X[2k] = FFT_{M/2}(x)[k] and X[2k+1] = FFT_{M/2}(x * W_M^n)[k].
template <int M, int NP, bool INV>
__device__ __forceinline__ void fft_pruned(const float2 (&in)[NP], float2 (&out)[M]) {
    static_assert(NP <= M / 2, "pruned FFT needs at least half of the input to be zero");
    float2 a[M / 2], b[M / 2];
    static_for<0, M / 2>([&](auto nc) {
        constexpr int n = decltype(nc)::value;
        if constexpr (n < NP) {
            a[n] = in[n];
            b[n] = twiddle<n, M, INV>(in[n]);
        } else {
            a[n] = make_float2(0.f, 0.f);
            b[n] = make_float2(0.f, 0.f);
        }
    });
    fft_radix2<M / 2, INV>(a);
    fft_radix2<M / 2, INV>(b);
    static_for<0, M / 2>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        out[2 * k] = a[k];
        out[2 * k + 1] = b[k];
    });
}

}  // namespace shwfs
