// sh_general.h -- SH forward for any subaperture sampling: np_sub up to 512, spectrum grid M a power of
// two from 16 to 1024, any number of CCD pixels per subaperture.  Included by shwfs.hip (uses NgsArgs,
// wave_sum and the phase reduction of that file).
//
// Same spectral formulation as the specialised kernels (B = P S Q^T for NGS, B = Re(P~ D Q~^T) for
// LGS, S = |FFT_M(E)|^2), evaluated with full FFTs on the M grid instead of register-resident
// pruned FFT_16/32 and lag-domain autocorrelations, so that the large subapertures of the
// reference's own set-ups (e.g. np 20 / M 64 LGS, np 80 / M 256, np 121 / M 256, np 240 / M 512,
// np 300 / M 1024; main/scao/params_eris.yml, params_morfeo_*.yml) run on the GPU.
//
// One work item = (frame, subaperture); a persistent grid of 256-thread blocks strides over the
// items.  Each block owns a slab of device scratch (allocated with the plan) for the intermediates
// that exceed LDS, and an LDS buffer in which CH vectors of length M are transformed at a time:
//   1  row pass     rows y < np of E = A exp(i psi) (zero padded to M)  -> FFT_M over x  -> G[y][kx]
//   2  column pass  columns kx of G (rows >= np are zero)              -> FFT_M over y  -> S = |.|^2
//        NGS: U[Y][kx] = sum_ky P[Y][ky] S[ky][kx]                      (npx x M, slab)
//        LGS: S kept (M x M real, slab)
//   3  NGS: B[Y][X] = sum_kx U[Y][kx] Q[X][kx]
//      LGS: FFT_M of S over kx (keep m_x = 0 .. np-1, the Hermitian half) -> T[ky][mx]
//           FFT_M of T over ky (keep |m_y| <= np-1), times the kernel spectrum -> D[my][mx]
//           V[Y][mx] = sum_my P~[Y][my] D[my][mx];  B[Y][X] = Re sum_mx c_mx V[Y][mx] Q~[X][mx]
//           (c_0 = 1, c_mx>0 = 2: D and the operators are transforms of real functions)
//   4  store the npx x npx tile scaled by the frame's Parseval scale, or its unscaled sum for the
//      non-Parseval rescale pass.
// A FoV mask that is not separable (round FoV mask clipping the cut window, sh.py:282) comes as K
// separable terms, B = sum_k P_k S Q_k^T (NGS) or Re sum_k P~_k D Q~_k^T (LGS): the NGS power
// spectrum S is then kept whole in the slab like the LGS one, and the K binnings accumulate into
// an npx x npx slab tile before the store.
// FFTs: four-step transforms (register FFT_R1 and FFT_R2 with M = R1 R2, one padded transpose
// through LDS), twiddles exp(-2 pi i k / M) from a table built on the host in FP64.  Accumulations of the binning products
// in FP32 over M terms (M <= 1024); the image sum of the non-Parseval mode in FP64.

struct GenArgs {
    const float2* tw;    // [M/2] exp(-2 pi i k / M)
    float* slab;         // [gridDim.x][slab_floats]
    int64_t slab_floats; // floats of scratch per block
    int M, log2M, np, ch, log2ch;   // ch = 1 << log2ch vectors per LDS chunk
    int terms;           // separable terms of the FoV mask (1: separable mask)
    int64_t term_stride; // floats per term in the row / column operators
};

constexpr int kGenThreads = 256;
constexpr int kGenRowsPerTask = 8;         // rows Y of V per task of the LGS product V = P~ D
// Complex working vectors per block: 32 KiB for M <= 256 (two blocks per CU with the transpose
// scratch), 64 KiB for M >= 512 (one block per CU; such plans have few, large subapertures, so
// the longer chunks matter more than the occupancy).  The smaller one bounds the LGS V tile.
constexpr int kGenLdsF2 = 4096 + 64;       // 16 rows of 256 + 1 (the LDS rows are padded to M + 1)
constexpr int kGenLdsBigF2 = 8192 + 128;   // 16 rows of 512 + 1, 8 of 1024 + 1
constexpr int gen_cap(int M) { return M >= 512 ? kGenLdsBigF2 : kGenLdsF2; }
// vectors per chunk: the largest power of two whose padded rows fit
inline int gen_chunk_log2(int M) {
    int l = 0;
    while ((2 << l) * (M + 1) <= gen_cap(M)) ++l;
    return l;
}

// Forward FFT (exp(-2 pi i jk/M), unnormalised) of `vecs` vectors of length M = R1 R2 held at
// buf[v * vs + i] (vs >= M: the callers pad rows to M + 1 so that their column-wise LDS accesses
// are odd-strided), in four steps (Cooley-Tukey with the transpose through LDS):
//   1  for each column n2 of the R1 x R2 view x[R2 n1 + n2]: FFT_R1 over n1 in registers, times
//      the twiddle W_M^(n2 k1), written to the scratch row n2 (rows padded to R1 + 1 complex)
//   2  for each k1: FFT_R2 over n2 of scratch column k1 in registers -> X[k1 + R1 k2], written
//      back to buf in natural order
// Every LDS access of a wave is to consecutive or odd-strided words (no bank conflicts), and a
// transform costs two barriers instead of the log2(M) + 1 of a radix-2 pass per stage.  The
// register FFTs are the compile-time ones of fft_reg.h.  Precondition: buf written and the
// block synchronised; the result is visible to the block on return.
template <int R1, int R2>
__device__ void gen_fft_rr(float2* buf, float2* scr, int vecs, int vs, const float2* __restrict__ tw) {
    constexpr int M = R1 * R2, RP = R1 + 1;
    for (int task = threadIdx.x; task < vecs * R2; task += blockDim.x) {
        const int v = task / R2, n2 = task - v * R2;
        const float2* p = buf + v * vs + n2;
        float2 x[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) x[n1] = p[n1 * R2];
        fft_radix2<R1, false>(x);
        float2* o = scr + (v * R2 + n2) * RP;
        o[0] = x[0];
#pragma unroll
        for (int k1 = 1; k1 < R1; ++k1) {
            const int e = n2 * k1;                       // < M; the table holds W^e for e < M/2
            float2 w = tw[e & (M / 2 - 1)];
            if (e & (M / 2)) w = make_float2(-w.x, -w.y);
            o[k1] = cmul(x[k1], w);
        }
    }
    __syncthreads();
    for (int task = threadIdx.x; task < vecs * R1; task += blockDim.x) {
        const int v = task / R1, k1 = task - v * R1;
        const float2* sc = scr + v * R2 * RP + k1;
        float2 x[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) x[n2] = sc[n2 * RP];
        fft_radix2<R2, false>(x);
        float2* o = buf + v * vs + k1;
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) o[k2 * R1] = x[k2];
    }
    __syncthreads();
}

// The factorisations of one kernel instantiation: M <= 256 (register transforms up to FFT_16, so
// the kernel keeps two waves per SIMD) or M = 512, 1024 (up to FFT_32).
template <int CAP>
__device__ void gen_fft(float2* buf, float2* scr, int vecs, int vs, int log2M, const float2* __restrict__ tw) {
    if constexpr (CAP == kGenLdsF2) {
        switch (log2M) {
            case 4: gen_fft_rr<4, 4>(buf, scr, vecs, vs, tw); break;
            case 5: gen_fft_rr<4, 8>(buf, scr, vecs, vs, tw); break;
            case 6: gen_fft_rr<8, 8>(buf, scr, vecs, vs, tw); break;
            case 7: gen_fft_rr<8, 16>(buf, scr, vecs, vs, tw); break;
            default: gen_fft_rr<16, 16>(buf, scr, vecs, vs, tw); break;  // log2M = 8
        }
    } else {
        if (log2M == 9) gen_fft_rr<16, 32>(buf, scr, vecs, vs, tw);
        else gen_fft_rr<32, 32>(buf, scr, vecs, vs, tw);                 // log2M = 10
    }
}

// Slab layout of one block (floats): NGS: G [np][M] complex, U [npx][M];
// LGS: G [np][M] complex (reused for T [M][np] complex), S [M][M], D [L][np] complex.
// With K > 1 mask terms: NGS G, S [M][M], U [npx][M], Bacc [npx][npx]; LGS as above + Bacc.
__device__ __forceinline__ void gen_slab(const GenArgs& q, int npx, bool lgs, float*& base, float2*& G, float*& U,
                                         float*& S, float2*& T, float2*& D, float*& Bacc) {
    base = q.slab + (int64_t)blockIdx.x * q.slab_floats;
    G = reinterpret_cast<float2*>(base);
    T = G;
    const int64_t g_f = 2LL * q.np * q.M;
    const int64_t t_f = 2LL * q.M * q.np;
    const int64_t gt = g_f > t_f ? g_f : t_f;
    const int64_t mm = (int64_t)q.M * q.M;
    S = base + gt;
    if (lgs) {
        U = base + gt;                                 // unused
        D = reinterpret_cast<float2*>(base + gt + mm);
        Bacc = base + gt + mm + 2LL * (2 * q.np - 1) * q.np;
    } else {
        U = q.terms > 1 ? base + gt + mm : base + gt;
        D = nullptr;
        Bacc = U + (int64_t)npx * q.M;
    }
}

template <int CAP>
__global__ void __launch_bounds__(kGenThreads)
sh_general_kernel(NgsArgs g, GenArgs q, int n_frames) {
    __shared__ float2 buf[CAP];
    __shared__ float2 fscr[CAP + CAP / 4];           // padded transpose scratch of the FFTs (R1 >= 4)
    __shared__ float2 tw[512];
    __shared__ double red[kGenThreads / 32];
    __shared__ int lit_flag;
    const int M = q.M, log2M = q.log2M, np = q.np, CH = q.ch;
    const int VS = M + 1;                            // LDS row stride of the working vectors (odd)
    const int lch = q.log2ch;
    const int npx = g.npx, n = g.n_lenses;
    const bool lgs = g.kspec != nullptr;
    const int L = 2 * np - 1;
    for (int k = threadIdx.x; k < (M >> 1); k += blockDim.x) tw[k] = q.tw[k];
    float* base;
    float2 *G, *T, *D;
    float *U, *S, *Bacc;
    gen_slab(q, npx, lgs, base, G, U, S, T, D, Bacc);
    const int K = q.terms;
    const int64_t n_items = (int64_t)n_frames * g.n_sub;
    const double k_rev = g.k * (1.0 / kTwoPi), t_rev = g.tilt * (1.0 / kTwoPi);
    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int f = (int)(item / g.n_sub);
        const int sub = (int)(item - (int64_t)f * g.n_sub);
        const int i = sub / n, j = sub - (sub / n) * n;
        const float* Af = g.A + (int64_t)f * g.a_stride;
        const float* Pf = g.phase + (int64_t)f * g.ph_stride;
        const int64_t off = (int64_t)(i * np) * g.ef_size + (int64_t)j * np;
        const float fs = g.parseval ? g.fscale[f * g.scale_fstep] : 1.f;
        float* ccdf = g.ccd + (int64_t)f * g.ccd_stride;
        const int W = n * npx;
        __syncthreads();                                   // slab and LDS of the previous item are free
        if (threadIdx.x == 0) lit_flag = 0;
        __syncthreads();
        bool lit = false;
        for (int e = threadIdx.x; e < np * np; e += blockDim.x) {
            const int y = e / np, x = e - (e / np) * np;
            lit |= Af[off + (int64_t)y * g.ef_size + x] != 0.f;
        }
        if (lit) lit_flag = 1;                             // benign race: every writer stores 1
        __syncthreads();
        if (!lit_flag) {                                   // zero image: 0 x scale, no transform
            const float v = g.parseval ? 0.f * fs : 0.f;
            for (int e = threadIdx.x; e < npx * npx; e += blockDim.x) {
                const int Y = e / npx, X = e - (e / npx) * npx;
                ccdf[(int64_t)(i * npx + Y) * W + j * npx + X] = v;
            }
            if (!g.parseval && threadIdx.x == 0) g.bpart[(int64_t)f * g.n_parts + sub] = 0.0;
            continue;
        }

        // 1: row transforms of the field (rows y < np), CH rows at a time
        for (int y0 = 0; y0 < np; y0 += CH) {
            const int rows = min(CH, np - y0);
            for (int e = threadIdx.x; e < rows * M; e += blockDim.x) {
                const int r = e >> log2M, x = e & (M - 1);
                float2 v = make_float2(0.f, 0.f);
                if (x < np) {
                    const int y = y0 + r;
                    const int64_t o = off + (int64_t)y * g.ef_size + x;
                    const float a = Af[o];
                    if (a != 0.f) {
                        double rv = fma((double)Pf[o], k_rev, -t_rev * (2.0 * g.a0 + (double)(x + y)));
                        rv -= rint(rv);
                        const float rf = (float)rv;
                        v = make_float2(a * __builtin_amdgcn_cosf(rf), a * __builtin_amdgcn_sinf(rf));
                    }
                }
                buf[r * VS + x] = v;
            }
            __syncthreads();
            gen_fft<CAP>(buf, fscr, rows, VS, log2M, tw);
            for (int e = threadIdx.x; e < rows * M; e += blockDim.x)
                G[(int64_t)y0 * M + e] = buf[(e >> log2M) * VS + (e & (M - 1))];
            __syncthreads();
        }

        // 2: column transforms, |.|^2, and the row binning (NGS) or the power spectrum (LGS)
        const int lc2 = lch < log2M ? lch : log2M;       // stage 2 chunks: 1 << lc2 columns
        for (int c0 = 0; c0 < M; c0 += CH) {
            const int cols = min(CH, M - c0);
            for (int e = threadIdx.x; e < cols * M; e += blockDim.x) {
                const int y = e >> lc2, c = e & (cols - 1);              // consecutive threads: consecutive G words
                buf[c * VS + y] = y < np ? G[(int64_t)y * M + c0 + c] : make_float2(0.f, 0.f);
            }
            __syncthreads();
            gen_fft<CAP>(buf, fscr, cols, VS, log2M, tw);
            for (int e = threadIdx.x; e < cols * M; e += blockDim.x) {
                float2* hp = buf + (e >> log2M) * VS + (e & (M - 1));
                const float2 h = *hp;
                *hp = make_float2(fmaf(h.x, h.x, h.y * h.y), 0.f);
            }
            __syncthreads();
            if (!lgs && K == 1) {
                // U[Y][kx] = sum_ky P[Y][ky] S[ky][kx] for the chunk's columns; P is staged in the
                // transpose scratch (free between transforms) when it fits, so that the tasks of a
                // chunk read it from LDS rather than each from L2
                const bool pstage = npx * M <= 2 * (CAP + CAP / 4);
                float* ps = reinterpret_cast<float*>(fscr);
                if (pstage) {
                    for (int e = threadIdx.x; e < npx * M; e += blockDim.x) ps[e] = g.roww[e];
                    __syncthreads();
                }
                for (int e = threadIdx.x; e < npx * cols; e += blockDim.x) {
                    const int Y = e / cols, c = e - (e / cols) * cols;
                    const float* pr = pstage ? ps + Y * M : g.roww + (int64_t)Y * M;
                    const float2* sc = buf + c * VS;
                    float acc = 0.f;
                    for (int ky = 0; ky < M; ++ky) acc = fmaf(pr[ky], sc[ky].x, acc);
                    U[(int64_t)Y * M + c0 + c] = acc;
                }
            } else {                                   // LGS, or NGS with K mask terms: keep S
                for (int e = threadIdx.x; e < cols * M; e += blockDim.x) {
                    const int ky = e >> lc2, c = e & (cols - 1);
                    S[(int64_t)ky * M + c0 + c] = buf[c * VS + ky].x;
                }
            }
            __syncthreads();
        }

        double bsum = 0.0;
        if (!lgs && K > 1) {
            // 3 (NGS, K mask terms): Bacc = sum_k (P_k S) Q_k^T
            for (int k = 0; k < K; ++k) {
                const float* Pk = g.roww + (int64_t)k * q.term_stride;
                const float* Qk = g.colw + (int64_t)k * q.term_stride;
                for (int e = threadIdx.x; e < npx * M; e += blockDim.x) {
                    const int Y = e >> log2M, kx = e & (M - 1);
                    const float* pr = Pk + (int64_t)Y * M;
                    float acc = 0.f;
                    for (int ky = 0; ky < M; ++ky) acc = fmaf(pr[ky], S[(int64_t)ky * M + kx], acc);
                    U[e] = acc;
                }
                __syncthreads();
                for (int e = threadIdx.x; e < npx * npx; e += blockDim.x) {
                    const int Y = e / npx, X = e - (e / npx) * npx;
                    const float* ur = U + (int64_t)Y * M;
                    const float* qr = Qk + (int64_t)X * M;
                    float acc = k == 0 ? 0.f : Bacc[e];
                    for (int kx = 0; kx < M; ++kx) acc = fmaf(ur[kx], qr[kx], acc);
                    Bacc[e] = acc;
                }
                __syncthreads();
            }
            for (int e = threadIdx.x; e < npx * npx; e += blockDim.x) {
                const int Y = e / npx, X = e - (e / npx) * npx;
                const float acc = Bacc[e];
                bsum += (double)acc;
                ccdf[(int64_t)(i * npx + Y) * W + j * npx + X] = g.parseval ? acc * fs : acc;
            }
        } else if (!lgs) {
            // 3 (NGS): B[Y][X] = sum_kx U[Y][kx] Q[X][kx], with U and Q staged in LDS when they fit
            const bool qstage = npx * M <= 2 * CAP;
            float* us = reinterpret_cast<float*>(buf);
            float* qs = reinterpret_cast<float*>(fscr);
            if (qstage) {
                for (int e = threadIdx.x; e < npx * M; e += blockDim.x) {
                    us[e] = U[e];
                    qs[e] = g.colw[e];
                }
                __syncthreads();
            }
            for (int e = threadIdx.x; e < npx * npx; e += blockDim.x) {
                const int Y = e / npx, X = e - (e / npx) * npx;
                const float* ur = qstage ? us + Y * M : U + (int64_t)Y * M;
                const float* qr = qstage ? qs + X * M : g.colw + (int64_t)X * M;
                float acc = 0.f;
                for (int kx = 0; kx < M; ++kx) acc = fmaf(ur[kx], qr[kx], acc);
                bsum += (double)acc;
                ccdf[(int64_t)(i * npx + Y) * W + j * npx + X] = g.parseval ? acc * fs : acc;
            }
        } else {
            // 3a (LGS): FFT of S over kx per row ky, keep m_x = 0 .. np-1  ->  T[ky][mx]
            for (int r0 = 0; r0 < M; r0 += CH) {
                const int rows = min(CH, M - r0);
                for (int e = threadIdx.x; e < rows * M; e += blockDim.x)
                    buf[(e >> log2M) * VS + (e & (M - 1))] = make_float2(S[(int64_t)r0 * M + e], 0.f);
                __syncthreads();
                gen_fft<CAP>(buf, fscr, rows, VS, log2M, tw);
                for (int e = threadIdx.x; e < rows * np; e += blockDim.x) {
                    const int r = e / np, mx = e - (e / np) * np;
                    T[(int64_t)(r0 + r) * np + mx] = buf[r * VS + mx];
                }
                __syncthreads();
            }
            // 3b: FFT of T over ky per column mx, keep |m_y| <= np-1, times the kernel spectrum
            const float2* ks = g.kspec + (int64_t)sub * L * L;
            for (int c0 = 0; c0 < np; c0 += CH) {
                const int cols = min(CH, np - c0);
                for (int e = threadIdx.x; e < (M << lch); e += blockDim.x) {
                    const int ky = e >> lch, c = e & (CH - 1);
                    if (c < cols) buf[c * VS + ky] = T[(int64_t)ky * np + c0 + c];
                }
                __syncthreads();
                gen_fft<CAP>(buf, fscr, cols, VS, log2M, tw);
                for (int e = threadIdx.x; e < (L << lch); e += blockDim.x) {
                    const int ly = e >> lch, c = e & (CH - 1);            // ly = my + np - 1
                    if (c >= cols) continue;
                    const int mx = c0 + c, my = ly - (np - 1);
                    const float2 d = buf[c * VS + ((my + M) & (M - 1))];
                    const float2 kk = ks[(int64_t)ly * L + (mx + np - 1)];
                    D[(int64_t)ly * np + mx] = make_float2(d.x * kk.x - d.y * kk.y, d.x * kk.y + d.y * kk.x);
                }
                __syncthreads();
            }
            // 3c: V[Y][mx] = sum_my P~_k[Y][my] D[my][mx] (complex, in LDS if it fits, else in S)
            // 3d: B += Re sum_mx c_mx V[Y][mx] Q~_k[X][mx], over the K mask terms
            // When P~_k fits the transpose scratch and a block has fewer lag columns than threads, P~_k
            // is staged there and each task (mx, yg) keeps up to kGenRowsPerTask rows Y = yg, yg + YG,
            // ... in registers, so D is read once per task instead of once per (Y, mx); the lag
            // columns m_x >= 0 of Q~_k follow in the same scratch for 3d.  Same sums in the same
            // order either way.
            float2* V = (npx * np <= CAP) ? buf : reinterpret_cast<float2*>(S);
            const int YG = max((npx + kGenRowsPerTask - 1) / kGenRowsPerTask, min(npx, kGenThreads / np));
            // staging pays when a task keeps more than one row (small subapertures: one row each)
            const bool staged = YG < npx && npx * L <= CAP + CAP / 4 && npx * np <= CAP;
            for (int k = 0; k < K; ++k) {
                const float2* Pt = reinterpret_cast<const float2*>(g.roww + (int64_t)k * q.term_stride);
                const float2* Qt = reinterpret_cast<const float2*>(g.colw + (int64_t)k * q.term_stride);
                if (staged) {
                    for (int e = threadIdx.x; e < npx * L; e += blockDim.x) fscr[e] = Pt[e];
                    __syncthreads();
                    for (int task = threadIdx.x; task < np * YG; task += blockDim.x) {
                        const int yg = task / np, mx = task - (task / np) * np;
                        float ar[kGenRowsPerTask], ai[kGenRowsPerTask];
#pragma unroll
                        for (int r = 0; r < kGenRowsPerTask; ++r) ar[r] = ai[r] = 0.f;
                        for (int ly = 0; ly < L; ++ly) {
                            const float2 d = D[(int64_t)ly * np + mx];
#pragma unroll
                            for (int r = 0; r < kGenRowsPerTask; ++r) {
                                const int Y = yg + r * YG;
                                if (Y < npx) {
                                    const float2 p = fscr[Y * L + ly];
                                    ar[r] = fmaf(p.x, d.x, fmaf(-p.y, d.y, ar[r]));
                                    ai[r] = fmaf(p.x, d.y, fmaf(p.y, d.x, ai[r]));
                                }
                            }
                        }
#pragma unroll
                        for (int r = 0; r < kGenRowsPerTask; ++r) {
                            const int Y = yg + r * YG;
                            if (Y < npx) V[Y * np + mx] = make_float2(ar[r], ai[r]);
                        }
                    }
                    __syncthreads();
                    for (int e = threadIdx.x; e < npx * np; e += blockDim.x) {
                        const int X = e / np, mx = e - (e / np) * np;
                        fscr[e] = Qt[(int64_t)X * L + (mx + np - 1)];
                    }
                } else {
                    for (int e = threadIdx.x; e < npx * np; e += blockDim.x) {
                        const int Y = e / np, mx = e - (e / np) * np;
                        float ar = 0.f, ai = 0.f;
                        for (int ly = 0; ly < L; ++ly) {
                            const float2 p = Pt[(int64_t)Y * L + ly];
                            const float2 d = D[(int64_t)ly * np + mx];
                            ar = fmaf(p.x, d.x, fmaf(-p.y, d.y, ar));
                            ai = fmaf(p.x, d.y, fmaf(p.y, d.x, ai));
                        }
                        V[e] = make_float2(ar, ai);
                    }
                }
                __syncthreads();
                for (int e = threadIdx.x; e < npx * npx; e += blockDim.x) {
                    const int Y = e / npx, X = e - (e / npx) * npx;
                    float acc = (K == 1 || k == 0) ? 0.f : Bacc[e];
                    for (int mx = 0; mx < np; ++mx) {
                        const float2 v = V[Y * np + mx];
                        const float2 qq = staged ? fscr[X * np + mx] : Qt[(int64_t)X * L + (mx + np - 1)];
                        const float re = fmaf(v.x, qq.x, -v.y * qq.y);
                        acc = fmaf(mx == 0 ? 1.f : 2.f, re, acc);
                    }
                    if (k + 1 < K) {
                        Bacc[e] = acc;
                    } else {
                        bsum += (double)acc;
                        ccdf[(int64_t)(i * npx + Y) * W + j * npx + X] = g.parseval ? acc * fs : acc;
                    }
                }
                __syncthreads();
            }
        }
        if (!g.parseval) {
            const double tot = block_sum(bsum, red);
            if (threadIdx.x == 0) g.bpart[(int64_t)f * g.n_parts + sub] = tot;
        }
    }
}
