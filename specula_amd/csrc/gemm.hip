// gemm.hip -- FP32 matrix product on the gfx950 matrix cores for the products next to the SH path:
// the modal DM projection phase = sign * cmd . ifunc (processing_objects/dm.py:72-86; batches of
// push-pull command vectors make it a dense GEMM) and the slope reconstruction / modal filtering
// slopes . recmat (processing_objects/slopec.py:92-104).
//
//   C[i][j] = alpha * sum_k A[i][k] op(B)[k][j] + beta * C[i][j] + bias[j]
//
// A is row-major m x k (lda), op(B) is B (row-major k x n, ldb) or B^T (B row-major n x k).
// Block tile 64 x 64, four waves of 32 x 32 (2 x 2 tiles of v_mfma_f32_16x16x4f32), K in chunks
// of 16 staged through padded LDS tiles; edges are zero-filled on load and masked on store.
// FP32 operands and FP32 accumulation (the matrix cores' FP32 product is exact per term).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/shwfs.h"

int shwfs_internal_set_error(int code, const char* msg);

namespace {

typedef float floatx4_g __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;
constexpr int kKc = 16;

template <bool TRANS_B>
__global__ void __launch_bounds__(256)
gemm_f32_kernel(int m, int n, int k, float alpha, const float* __restrict__ A, int64_t lda,
                const float* __restrict__ B, int64_t ldb, float beta, const float* __restrict__ bias,
                float* __restrict__ C, int64_t ldc) {
    __shared__ float As[kTile][kKc + 1];
    __shared__ float Bs[kKc][kTile + 1];
    const int t = threadIdx.x;
    const int lane = t & 63, wv = t >> 6;
    const int i0 = blockIdx.y * kTile, j0 = blockIdx.x * kTile;
    const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
    floatx4_g acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = floatx4_g{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < k; k0 += kKc) {
        // A tile: 64 rows x 16 k, 4 consecutive k per thread
        {
            const int r = t >> 2, kq = (t & 3) * 4;
            const int gi = i0 + r;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int gk = k0 + kq + u;
                As[r][kq + u] = (gi < m && gk < k) ? A[(int64_t)gi * lda + gk] : 0.f;
            }
        }
        // op(B) tile: 16 k x 64 columns
        if constexpr (!TRANS_B) {
            const int r = t >> 4, cq = (t & 15) * 4;
            const int gk = k0 + r;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int gj = j0 + cq + u;
                Bs[r][cq + u] = (gk < k && gj < n) ? B[(int64_t)gk * ldb + gj] : 0.f;
            }
        } else {
            const int c = t >> 2, kq = (t & 3) * 4;
            const int gj = j0 + c;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int gk = k0 + kq + u;
                Bs[kq + u][c] = (gj < n && gk < k) ? B[(int64_t)gj * ldb + gk] : 0.f;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kKc; kk += 4) {
            const int kr = kk + (lane >> 4);
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const float av = As[wr + a * 16 + (lane & 15)][kr];
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const float bv = Bs[kr][wc + b * 16 + (lane & 15)];
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[a][b], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // accumulator layout: lane l, register r -> row 4 (l / 16) + r, column l % 16 of a 16 x 16 tile
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int gj = j0 + wc + b * 16 + (lane & 15);
            if (gj >= n) continue;
            const float bj = bias ? bias[gj] : 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = i0 + wr + a * 16 + 4 * (lane >> 4) + r;
                if (gi >= m) continue;
                float* c = C + (int64_t)gi * ldc + gj;
                float v = alpha * acc[a][b][r] + bj;
                if (beta != 0.f) v = fmaf(beta, *c, v);
                *c = v;
            }
        }
}

}  // namespace

extern "C" {

int shgemm_f32(int m, int n, int k, float alpha, const float* A, int64_t lda, const float* B, int64_t ldb,
               int trans_b, float beta, const float* bias, float* C, int64_t ldc, void* stream) {
    if (!A || !B || !C) return shwfs_internal_set_error(SHWFS_ERR_ARG, "null argument");
    if (m <= 0 || n <= 0 || k <= 0) return shwfs_internal_set_error(SHWFS_ERR_ARG, "non-positive matrix size");
    if (lda < k || ldc < n || ldb < (trans_b ? k : n))
        return shwfs_internal_set_error(SHWFS_ERR_ARG, "leading dimension smaller than the row length");
    if ((int64_t)m > 65535LL * kTile) return shwfs_internal_set_error(SHWFS_ERR_ARG, "m too large for one launch");
    const dim3 grid((n + kTile - 1) / kTile, (m + kTile - 1) / kTile);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (trans_b)
        hipLaunchKernelGGL(gemm_f32_kernel<true>, grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, bias, C,
                           ldc);
    else
        hipLaunchKernelGGL(gemm_f32_kernel<false>, grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, bias,
                           C, ldc);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return shwfs_internal_set_error(SHWFS_ERR_HIP, (std::string("kernel launch: ") + hipGetErrorString(e)).c_str());
    return SHWFS_OK;
}

}  // extern "C"
