// k_gemm.hip -- dense FP64 GEMM in the canonical order on the VALU.
//
// MI355X runs FP64 FMA on the vector ALU at the same 78.6 TFLOP/s as the FP64
// matrix cores, so a register-tiled v_fma_f64 GEMM loses nothing against
// MFMA while keeping one sequential fma chain per output entry (k ascending),
// the order the oracle reproduces bit for bit.
//
// Tiling: a 256-thread workgroup computes a 64x64 tile of C; each thread a
// 4x4 micro-tile with rows ty+16r and columns tx+16c (strided, so the LDS
// reads of a wave are broadcasts or distinct banks).  K advances in steps of
// 32 through LDS tiles As[k][i] and Bs[k][j]; the next step's global values
// are loaded into registers while the current step computes, so one
// workgroup per CU still overlaps HBM/L2 latency with FMAs.  Out-of-range
// entries load as 0, which leaves every fma chain unchanged.
#include "fh_internal.hpp"

namespace fh {

constexpr int GT = 64;  // tile edge
constexpr int GK = 32;  // k step
constexpr int GPT = GT * GK / 256;  // elements of each operand tile per thread (8)

__device__ __forceinline__ void gemm_load(const double* __restrict__ P, int ld, int trans, int rows, int kdim,
                                          int r0, int k0, int t, double (&reg)[GPT]) {
  // tile element e = t + 256*q; non-transposed operand: fastest index is the
  // row (contiguous in column-major), transposed: fastest is k
#pragma unroll
  for (int q = 0; q < GPT; ++q) {
    const int e = t + 256 * q;
    int rr, kk;
    if (!trans) {
      rr = e & (GT - 1);
      kk = e >> 6;
    } else {
      kk = e & (GK - 1);
      rr = e >> 5;
    }
    const int gr = r0 + rr, gk = k0 + kk;
    reg[q] = (gr < rows && gk < kdim) ? (trans ? P[(size_t)gr * ld + gk] : P[(size_t)gk * ld + gr]) : 0.0;
  }
}

__device__ __forceinline__ void gemm_store(double (*S)[GT + 1], int trans, int t, const double (&reg)[GPT]) {
#pragma unroll
  for (int q = 0; q < GPT; ++q) {
    const int e = t + 256 * q;
    int rr, kk;
    if (!trans) {
      rr = e & (GT - 1);
      kk = e >> 6;
    } else {
      kk = e & (GK - 1);
      rr = e >> 5;
    }
    S[kk][rr] = reg[q];
  }
}

// op(A) is m x k: element (i, kk) = transA ? A[i*lda + kk] : A[kk*lda + i]
// op(B) is k x n: element (kk, j) = transB ? B[kk*ldb + j] : B[j*ldb + kk]
// B is loaded as the "row" operand of its transpose: row j, column kk.
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
  double ra[GPT], rb[GPT];
  // op(B)^T (n x k) row j, col kk = op(B)(kk, j): stored "transposed" iff !transB
  gemm_load(A, lda, transA, m, k, i0, 0, t, ra);
  gemm_load(B, ldb, !transB, n, k, j0, 0, t, rb);
  for (int k0 = 0; k0 < k; k0 += GK) {
    gemm_store(As, transA, t, ra);
    gemm_store(Bs, !transB, t, rb);
    __syncthreads();
    if (k0 + GK < k) {
      gemm_load(A, lda, transA, m, k, i0, k0 + GK, t, ra);
      gemm_load(B, ldb, !transB, n, k, j0, k0 + GK, t, rb);
    }
#pragma unroll 8
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
    const int gi = i0 + ty + 16 * r;
    if (gi >= m) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int gj = j0 + tx + 16 * c;
      if (gj >= n) continue;
      const double tt = alpha * acc[r][c];
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
