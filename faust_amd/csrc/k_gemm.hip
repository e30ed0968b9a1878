// k_gemm.hip -- dense FP64 GEMM in the canonical order on the VALU.
//
// MI355X runs FP64 FMA on the vector ALU at the same 78.6 TFLOP/s as the FP64
// matrix cores, so a register-tiled v_fma_f64 GEMM loses nothing against
// MFMA while keeping one sequential fma chain per output entry (k ascending),
// the order the oracle reproduces bit for bit.
//
// Tiling: a 256-thread workgroup computes a 64x64 tile of C; each thread a
// 4x4 micro-tile with rows ty+16r and columns tx+16c (strided, so the LDS
// reads of one wave hit distinct banks).  K advances in steps of 16 through
// LDS tiles As[k][i] and Bs[k][j]; out-of-range entries load as 0, which
// leaves every fma chain unchanged.
#include "fh_internal.hpp"

namespace fh {

constexpr int GT = 64;  // tile edge
constexpr int GK = 16;  // k step

__global__ __launch_bounds__(256) void k_gemm_f64(int m, int n, int k, double alpha, const double* __restrict__ A,
                                                   int lda, int transA, const double* __restrict__ B, int ldb,
                                                   int transB, double beta, double* __restrict__ C, int ldc) {
  __shared__ double As[GK][GT + 1];
  __shared__ double Bs[GK][GT + 1];
  const int t = threadIdx.x;
  const int tx = t & 15, ty = t >> 4;
  const int i0 = blockIdx.x * GT, j0 = blockIdx.y * GT;
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;

  for (int k0 = 0; k0 < k; k0 += GK) {
    // op(A) tile: rows i0..i0+63, cols k0..k0+15 -> As[kk][ii]
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int e = t + 256 * q;  // 0..1023
      int ii, kk;
      if (!transA) {
        ii = e & 63;
        kk = e >> 6;
      } else {
        kk = e & 15;
        ii = e >> 4;
      }
      int gi = i0 + ii, gk = k0 + kk;
      double val = 0.0;
      if (gi < m && gk < k) val = transA ? A[(size_t)gi * lda + gk] : A[(size_t)gk * lda + gi];
      As[kk][ii] = val;
    }
    // op(B) tile: rows k0..k0+15, cols j0..j0+63 -> Bs[kk][jj]
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int e = t + 256 * q;
      int jj, kk;
      if (!transB) {
        kk = e & 15;
        jj = e >> 4;
      } else {
        jj = e & 63;
        kk = e >> 6;
      }
      int gj = j0 + jj, gk = k0 + kk;
      double val = 0.0;
      if (gj < n && gk < k) val = transB ? B[(size_t)gk * ldb + gj] : B[(size_t)gj * ldb + gk];
      Bs[kk][jj] = val;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = As[kk][ty + 16 * r];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = Bs[kk][tx + 16 * c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = fma(a[r], b[c], acc[r][c]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    int gi = i0 + ty + 16 * r;
    if (gi >= m) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      int gj = j0 + tx + 16 * c;
      if (gj >= n) continue;
      double tt = alpha * acc[r][c];
      double* cp = C + (size_t)gj * ldc + gi;
      *cp = (beta == 0.0) ? tt : fma(beta, *cp, tt);
    }
  }
}

void gemm_f64(char opA, char opB, int m, int n, int k, double alpha, const double* A, int lda, const double* B,
              int ldb, double beta, double* C, int ldc, hipStream_t st) {
  if (m <= 0 || n <= 0) return;
  dim3 grid((m + GT - 1) / GT, (n + GT - 1) / GT);
  hipLaunchKernelGGL(k_gemm_f64, grid, dim3(256), 0, st, m, n, k, alpha, A, lda, opA != 'N' ? 1 : 0, B, ldb,
                     opB != 'N' ? 1 : 0, beta, C, ldc);
  FH_LAUNCH_CHECK();
}

}  // namespace fh
