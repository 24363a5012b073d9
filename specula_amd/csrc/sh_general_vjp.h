// sh_general_vjp.h -- vector-Jacobian product of the any-size SH path (sh_general.h), NGS and LGS,
// any np_sub up to 512 and M up to 1024, separable or K-term FoV masks.  Included by shwfs.hip
// after VjpArgs (the dL/dI -> dL/dphi contract of sh_ngs_vjp_kernel: G_B = c (dL/dI - alpha) on
// each subaperture tile, c the frame's flux scale, alpha the cotangent offset of a self-normalised
// image).
//
// The image of one subaperture is a real-linear function of the M x M power spectrum
// S = |F|^2, F = FFT_M(E) (E = A exp(i psi) zero padded to M), so with G_S = dL/dS:
//   X = G_S conj(F),  H = FFT_M(X) on the np x np support,  dL/dphi = -2 k Im(H E)
// (k = 2 pi / lambda in rad per nm), and G_S is the transpose of the binning:
//   NGS:  G_S = sum_k P_k^T G_B Q_k                                   (B = sum_k P_k S Q_k^T)
//   LGS:  G_S = Re FFT2_M(A'),  A'[m] = c_mx K[m] sum_k (P~_k^T G_B Q~_k)[m]   on the lag grid
//         |m_y| <= np-1, 0 <= m_x <= np-1 (c_0 = 1, c_mx>0 = 2), placed at (m_y mod M, m_x)
//         (B = Re sum_k P~_k D Q~_k^T with D[m] = FFT_M(S)[m mod M] K[m], P~ carrying N^2/M^2).
// One work item = (frame, subaperture); a persistent grid of 256-thread blocks strides over the
// items; each block owns a slab of device scratch (allocated with the first VJP of a plan).
//
// Slab layout of one block (floats; each part starts on a 16-byte boundary):
//   Es [np][np] complex | F [M][M] complex | Gb [npx][npx] | R
//   R: NGS  Z_k [K][npx][M] (Z_k = G_B Q_k^T), then GS [M][M] when G_S is accumulated in parts
//      LGS  W1 [npx][np] complex, Ar [L][M] complex   (W1 = G_B Q~_k, Ar rows m_y of A')
// The NGS product G_S = sum_k P_k^T Z_k runs from LDS: chunks of gen_vjp_rows(M) rows of P_k and
// Z_k, each thread accumulating 4 x 4 blocks of G_S in registers.  One chunk and one term (every
// separable mask with npx <= gen_vjp_rows(M)): X is formed from the registers; otherwise the
// partial sums go through GS.  The LGS products stage P~_k and W1 in LDS when they fit.

struct GenVjpArgs {
    const float2* tw;    // [M/2] exp(-2 pi i k / M)
    float* slab;         // [gridDim.x][slab_floats]
    int64_t slab_floats;
    int M, log2M, np, ch, log2ch;
    int terms;
    int64_t term_stride; // floats per mask term in the row / column operators
    const float2* kspec; // LGS: [n_sub][L][L] kernel spectra on the lag grid, else null
};

// rows of P_k / Z_k per LDS chunk of the NGS product (the LDS vectors hold 2 CAP floats)
constexpr int gen_vjp_rows(int M) { return 2 * gen_cap(M) / M; }
constexpr bool gen_vjp_direct(int M, int npx, int terms) { return terms == 1 && npx <= gen_vjp_rows(M); }

struct GenVjpSlab {
    int64_t f, gb, r, gs, total;   // float offsets of F, Gb, R and GS; floats per block
};
__host__ __device__ inline GenVjpSlab gen_vjp_slab(int np, int M, int npx, bool lgs, int terms) {
    auto a4 = [](int64_t v) { return (v + 3) & ~(int64_t)3; };
    GenVjpSlab s;
    s.f = a4(2LL * np * np);
    s.gb = s.f + 2LL * M * M;
    s.r = a4(s.gb + (int64_t)npx * npx);
    if (lgs) {
        s.gs = s.r;
        s.total = s.r + 2LL * npx * np + 2LL * (2 * np - 1) * M;
    } else {
        s.gs = a4(s.r + (int64_t)terms * npx * M);
        s.total = s.gs + (gen_vjp_direct(M, npx, terms) ? 0 : (int64_t)M * M);
    }
    s.total = (s.total + 63) / 64 * 64;           // 256-byte aligned slabs
    return s;
}

template <int CAP>
__global__ void __launch_bounds__(kGenThreads)
sh_general_vjp_kernel(VjpArgs g, GenVjpArgs q, int n_frames) {
    __shared__ __align__(16) float2 buf[CAP];
    __shared__ __align__(16) float2 fscr[CAP + CAP / 4];
    __shared__ float2 tw[512];
    __shared__ int lit_flag;
    const int M = q.M, log2M = q.log2M, np = q.np, CH = q.ch, lch = q.log2ch;
    const int VS = M + 1;
    const int npx = g.npx, n = g.n_lenses;
    const bool lgs = q.kspec != nullptr;
    const int L = 2 * np - 1, K = q.terms;
    for (int k = threadIdx.x; k < (M >> 1); k += blockDim.x) tw[k] = q.tw[k];
    const GenVjpSlab so = gen_vjp_slab(np, M, npx, lgs, K);
    float* base = q.slab + (int64_t)blockIdx.x * q.slab_floats;
    float2* Es = reinterpret_cast<float2*>(base);
    float2* F = reinterpret_cast<float2*>(base + so.f);
    float* Gb = base + so.gb;
    float* R = base + so.r;
    float* Z = R;                                             // NGS
    float* GS = base + so.gs;                                 // NGS, parts of G_S
    float2* W1 = reinterpret_cast<float2*>(R);                // LGS
    float2* Ar = W1 + (int64_t)npx * np;                      // LGS
    const int64_t n_items = (int64_t)n_frames * g.n_sub;
    const double k_rev = g.k * (1.0 / kTwoPi), t_rev = g.tilt * (1.0 / kTwoPi);
    const int lc2 = lch < log2M ? lch : log2M;
    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int f = (int)(item / g.n_sub);
        const int sub = (int)(item - (int64_t)f * g.n_sub);
        const int i = sub / n, j = sub - (sub / n) * n;
        const float* Af = g.A + (int64_t)f * g.a_stride;
        const float* Pf = g.phase + (int64_t)f * g.ph_stride;
        const int64_t off = (int64_t)(i * np) * g.ef_size + (int64_t)j * np;
        float* gph = g.gphase + (int64_t)f * g.gph_stride + off;
        float c, alpha = 0.f;
        if (g.cg) {
            c = (float)g.cg[f].x;
            alpha = (float)g.cg[f].y;
        } else {
            const double* am = g.amp + (int64_t)f * g.amp_fstep;
            c = (float)(g.S0 * g.pixel_area * am[0] / (g.norm_const * am[1]));
        }
        __syncthreads();                                      // slab and LDS of the previous item are free
        if (threadIdx.x == 0) lit_flag = 0;
        __syncthreads();
        // 0: the field of the subaperture (kept for the last step); a dark one has a zero gradient
        bool lit = false;
        for (int e = threadIdx.x; e < np * np; e += blockDim.x) {
            const int y = e / np, x = e - (e / np) * np;
            const int64_t o = (int64_t)y * g.ef_size + x;
            const float a = Af[off + o];
            float2 v = make_float2(0.f, 0.f);
            if (a != 0.f) {
                lit = true;
                double rv = fma((double)Pf[off + o], k_rev, -t_rev * (2.0 * g.a0 + (double)(x + y)));
                rv -= rint(rv);
                const float rf = (float)rv;
                v = make_float2(a * __builtin_amdgcn_cosf(rf), a * __builtin_amdgcn_sinf(rf));
            }
            Es[e] = v;
        }
        if (lit) lit_flag = 1;
        __syncthreads();
        if (!lit_flag) {
            for (int e = threadIdx.x; e < np * np; e += blockDim.x) {
                const int y = e / np, x = e - (e / np) * np;
                gph[(int64_t)y * g.ef_size + x] = 0.f;
            }
            continue;
        }
        // 1: F = FFT_M(E): rows (y < np) into F, then columns in place
        for (int y0 = 0; y0 < np; y0 += CH) {
            const int rows = min(CH, np - y0);
            for (int e = threadIdx.x; e < rows * M; e += blockDim.x) {
                const int r = e >> log2M, x = e & (M - 1);
                buf[r * VS + x] = x < np ? Es[(int64_t)(y0 + r) * np + x] : make_float2(0.f, 0.f);
            }
            __syncthreads();
            gen_fft<CAP>(buf, fscr, rows, VS, log2M, tw);
            for (int e = threadIdx.x; e < rows * M; e += blockDim.x)
                F[(int64_t)y0 * M + e] = buf[(e >> log2M) * VS + (e & (M - 1))];
            __syncthreads();
        }
        for (int c0 = 0; c0 < M; c0 += CH) {
            const int cols = min(CH, M - c0);
            for (int e = threadIdx.x; e < cols * M; e += blockDim.x) {
                const int y = e >> lc2, c = e & (cols - 1);
                buf[c * VS + y] = y < np ? F[(int64_t)y * M + c0 + c] : make_float2(0.f, 0.f);
            }
            __syncthreads();
            gen_fft<CAP>(buf, fscr, cols, VS, log2M, tw);
            for (int e = threadIdx.x; e < cols * M; e += blockDim.x) {
                const int ky = e >> lc2, c = e & (cols - 1);
                F[(int64_t)ky * M + c0 + c] = buf[c * VS + ky];
            }
            __syncthreads();
        }
        // 2: G_B = c (dL/dI - alpha) on the subaperture tile
        for (int e = threadIdx.x; e < npx * npx; e += blockDim.x) {
            const int Y = e / npx, X = e - (e / npx) * npx;
            Gb[e] = (g.gccd[f * g.gccd_stride + (int64_t)(i * npx + Y) * (n * npx) + j * npx + X] - alpha) * c;
        }
        __syncthreads();
        if (!lgs) {
            // 3 (NGS): Z_k = G_B Q_k^T, then X = (sum_k P_k^T Z_k) conj(F) in place of F
            for (int k = 0; k < K; ++k) {
                const float* Qk = g.colw + (int64_t)k * q.term_stride;
                float* Zk = Z + (int64_t)k * npx * M;
                for (int e = threadIdx.x; e < npx * M; e += blockDim.x) {
                    const int Y = e >> log2M, kx = e & (M - 1);
                    float acc = 0.f;
                    for (int X = 0; X < npx; ++X) acc = fmaf(Gb[Y * npx + X], Qk[(int64_t)X * M + kx], acc);
                    Zk[e] = acc;
                }
            }
            __syncthreads();
            // G_S[ky][kx] = sum_k sum_Y P_k[Y][ky] Z_k[Y][kx] from LDS chunks of rows of P_k and Z_k,
            // a 4 x 4 block of G_S per thread and step
            float* Pl = reinterpret_cast<float*>(buf);
            float* Zl = reinterpret_cast<float*>(fscr);
            const int yc = gen_vjp_rows(M);
            const int lm4 = log2M - 2, nb = 1 << (2 * lm4);
            for (int k = 0; k < K; ++k) {
                const float* Pk = g.roww + (int64_t)k * q.term_stride;
                const float* Zk = Z + (int64_t)k * npx * M;
                for (int y0 = 0; y0 < npx; y0 += yc) {
                    const int rows = min(yc, npx - y0);
                    for (int e = threadIdx.x; e < rows * M; e += blockDim.x) {
                        Pl[e] = Pk[(int64_t)y0 * M + e];
                        Zl[e] = Zk[(int64_t)y0 * M + e];
                    }
                    __syncthreads();
                    const bool first = k == 0 && y0 == 0, last = k == K - 1 && y0 + rows >= npx;
                    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
                        const int ky0 = (b >> lm4) << 2, kx0 = (b & ((1 << lm4) - 1)) << 2;
                        float acc[4][4];
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                            if (!first) v = *reinterpret_cast<const float4*>(GS + (int64_t)(ky0 + r) * M + kx0);
                            acc[r][0] = v.x; acc[r][1] = v.y; acc[r][2] = v.z; acc[r][3] = v.w;
                        }
                        for (int Y = 0; Y < rows; ++Y) {
                            const float4 p4 = *reinterpret_cast<const float4*>(Pl + Y * M + ky0);
                            const float4 z4 = *reinterpret_cast<const float4*>(Zl + Y * M + kx0);
                            const float pv[4] = {p4.x, p4.y, p4.z, p4.w}, zv[4] = {z4.x, z4.y, z4.z, z4.w};
#pragma unroll
                            for (int r = 0; r < 4; ++r)
#pragma unroll
                                for (int c2 = 0; c2 < 4; ++c2) acc[r][c2] = fmaf(pv[r], zv[c2], acc[r][c2]);
                        }
                        if (last) {                                   // X = G_S conj(F) in place of F
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                float2* fp = F + (int64_t)(ky0 + r) * M + kx0;
#pragma unroll
                                for (int c2 = 0; c2 < 4; ++c2) {
                                    const float2 fv = fp[c2];
                                    fp[c2] = make_float2(acc[r][c2] * fv.x, -acc[r][c2] * fv.y);
                                }
                            }
                        } else {
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                *reinterpret_cast<float4*>(GS + (int64_t)(ky0 + r) * M + kx0) =
                                    make_float4(acc[r][0], acc[r][1], acc[r][2], acc[r][3]);
                        }
                    }
                    __syncthreads();
                }
            }
        } else {
            // 3 (LGS): A' on the lag grid (rows m_y, columns m_x < np, zero padded to M), summed over
            // the mask terms, times the kernel spectrum; Ar = its FFT over m_x; then the FFT over m_y
            // gives G_S, and X = G_S conj(F) in place of F
            const float2* ks = q.kspec + (int64_t)sub * L * L;
            for (int e = threadIdx.x; e < L * M; e += blockDim.x) Ar[e] = make_float2(0.f, 0.f);
            __syncthreads();
            // P~_k and W1 in LDS when they fit (buf and fscr are free until the transforms below)
            const bool p_lds = npx * L <= CAP, w_lds = npx * np <= CAP + CAP / 4;
            float2* Wv = w_lds ? fscr : W1;
            for (int k = 0; k < K; ++k) {
                const float2* Pg = reinterpret_cast<const float2*>(g.roww + (int64_t)k * q.term_stride);
                const float2* Qt = reinterpret_cast<const float2*>(g.colw + (int64_t)k * q.term_stride);
                const float2* Pt = Pg;
                if (p_lds) {
                    for (int e = threadIdx.x; e < npx * L; e += blockDim.x) buf[e] = Pg[e];
                    Pt = buf;
                }
                for (int e = threadIdx.x; e < npx * np; e += blockDim.x) {     // W1 = G_B Q~ (lag columns m_x)
                    const int Y = e / np, mx = e - (e / np) * np;
                    float ar = 0.f, ai = 0.f;
                    for (int X = 0; X < npx; ++X) {
                        const float gb = Gb[Y * npx + X];
                        const float2 qq = Qt[(int64_t)X * L + (mx + np - 1)];
                        ar = fmaf(gb, qq.x, ar);
                        ai = fmaf(gb, qq.y, ai);
                    }
                    Wv[e] = make_float2(ar, ai);
                }
                __syncthreads();
                for (int e = threadIdx.x; e < L * np; e += blockDim.x) {       // A' += c_mx P~^T W1
                    const int ly = e / np, mx = e - (e / np) * np;
                    float ar = 0.f, ai = 0.f;
                    for (int Y = 0; Y < npx; ++Y) {
                        const float2 p = Pt[Y * L + ly];
                        const float2 w = Wv[Y * np + mx];
                        ar = fmaf(p.x, w.x, fmaf(-p.y, w.y, ar));
                        ai = fmaf(p.x, w.y, fmaf(p.y, w.x, ai));
                    }
                    const float cm = mx == 0 ? 1.f : 2.f;
                    float2* a = Ar + (int64_t)ly * M + mx;
                    *a = make_float2(a->x + cm * ar, a->y + cm * ai);
                }
                __syncthreads();
            }
            for (int e = threadIdx.x; e < L * np; e += blockDim.x) {           // times K[m]
                const int ly = e / np, mx = e - (e / np) * np;
                const float2 kk = ks[(int64_t)ly * L + (mx + np - 1)];
                float2* a = Ar + (int64_t)ly * M + mx;
                *a = make_float2(a->x * kk.x - a->y * kk.y, a->x * kk.y + a->y * kk.x);
            }
            __syncthreads();
            for (int r0 = 0; r0 < L; r0 += CH) {                               // FFT over m_x of rows m_y
                const int rows = min(CH, L - r0);
                for (int e = threadIdx.x; e < rows * M; e += blockDim.x)
                    buf[(e >> log2M) * VS + (e & (M - 1))] = Ar[(int64_t)r0 * M + e];
                __syncthreads();
                gen_fft<CAP>(buf, fscr, rows, VS, log2M, tw);
                for (int e = threadIdx.x; e < rows * M; e += blockDim.x)
                    Ar[(int64_t)r0 * M + e] = buf[(e >> log2M) * VS + (e & (M - 1))];
                __syncthreads();
            }
            for (int c0 = 0; c0 < M; c0 += CH) {                               // FFT over m_y, X = G_S conj(F)
                const int cols = min(CH, M - c0);
                for (int e = threadIdx.x; e < cols * M; e += blockDim.x) {
                    const int y = e >> lc2, c = e & (cols - 1);               // y: grid row m_y mod M
                    const int ly = y < np ? y + np - 1 : (y >= M - (np - 1) ? y - M + np - 1 : -1);
                    buf[c * VS + y] = ly >= 0 ? Ar[(int64_t)ly * M + c0 + c] : make_float2(0.f, 0.f);
                }
                __syncthreads();
                gen_fft<CAP>(buf, fscr, cols, VS, log2M, tw);
                for (int e = threadIdx.x; e < cols * M; e += blockDim.x) {
                    const int ky = e >> lc2, c = e & (cols - 1);
                    const float gs = buf[c * VS + ky].x;
                    float2* fp = F + (int64_t)ky * M + c0 + c;
                    const float2 fv = *fp;
                    *fp = make_float2(gs * fv.x, -gs * fv.y);
                }
                __syncthreads();
            }
        }
        // 4: H = FFT_M(X) on the np x np support: columns over ky (keep rows < np), then rows
        for (int c0 = 0; c0 < M; c0 += CH) {
            const int cols = min(CH, M - c0);
            for (int e = threadIdx.x; e < cols * M; e += blockDim.x) {
                const int ky = e >> lc2, c = e & (cols - 1);
                buf[c * VS + ky] = F[(int64_t)ky * M + c0 + c];
            }
            __syncthreads();
            gen_fft<CAP>(buf, fscr, cols, VS, log2M, tw);
            for (int e = threadIdx.x; e < cols * np; e += blockDim.x) {
                const int ny = e / cols, c = e - (e / cols) * cols;
                F[(int64_t)ny * M + c0 + c] = buf[c * VS + ny];
            }
            __syncthreads();
        }
        const float two_k = (float)(2.0 * g.k);
        for (int y0 = 0; y0 < np; y0 += CH) {
            const int rows = min(CH, np - y0);
            for (int e = threadIdx.x; e < rows * M; e += blockDim.x)
                buf[(e >> log2M) * VS + (e & (M - 1))] = F[(int64_t)y0 * M + e];
            __syncthreads();
            gen_fft<CAP>(buf, fscr, rows, VS, log2M, tw);
            for (int e = threadIdx.x; e < rows * np; e += blockDim.x) {
                const int r = e / np, x = e - (e / np) * np;
                const float2 h = buf[r * VS + x];
                const float2 ev = Es[(int64_t)(y0 + r) * np + x];
                gph[(int64_t)(y0 + r) * g.ef_size + x] = -two_k * (h.x * ev.y + h.y * ev.x);   // Im(H E)
            }
            __syncthreads();
        }
    }
}
