// k_spmm.hip -- FP64 sparse x dense and dense x sparse products in the
// canonical order (one fma chain per output entry, k ascending).
//
// Layout: dense operands column-major with a leading dimension.  A wave holds
// 64 consecutive output rows of one output column, so the Y stores and (for
// dense x sparse) the A loads are coalesced; X gathers X[colind, j] stay in
// one column.  op(S) = S^T walks the CSC twin of S, whose rows are ascending
// inside each column, so the k order is the same as the oracle's.
#include "fh_internal.hpp"

namespace fh {

// Y(i,j) = alpha * sum_p v[p] * X(ci[p], j) + beta * Y(i,j)
// transX: X(k,j) is read as X[j + k*ldx] (op(X) = X^T).
template <bool TRANSX>
__global__ __launch_bounds__(256) void k_csr_spmm_f64(int nrows, int ncols, const int* __restrict__ rp,
                                                       const int* __restrict__ ci, const double* __restrict__ v,
                                                       const double* __restrict__ X, int ldx,
                                                       double* __restrict__ Y, int ldy, double alpha, double beta) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= nrows || j >= ncols) return;
  const int b = rp[i], e = rp[i + 1];
  double acc = 0.0;
  // batches of 8 terms: the index/value loads and the gathers are issued
  // unconditionally (indices clamped inside the row) before the fma chain
  // runs in k order; terms past the row end are not accumulated
  const double* xj = X + (size_t)j * ldx;
  for (int base = b; base < e; base += 8) {
    const int nb = min(8, e - base);
    int cc[8];
    double ww[8], xx[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int p = base + min(q, nb - 1);
      cc[q] = ci[p];
      ww[q] = v[p];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) xx[q] = TRANSX ? X[(size_t)cc[q] * ldx + j] : xj[cc[q]];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double f = fma(ww[q], xx[q], acc);
      acc = q < nb ? f : acc;
    }
  }
  const double t = alpha * acc;
  double* y = Y + (size_t)j * ldy + i;
  *y = (beta == 0.0) ? t : fma(beta, *y, t);
}

void csr_spmm_f64_x(const DMat& S, char op, const double* X, int ldx, int transX, int ncols, double* Y, int ldy,
                    double alpha, double beta, hipStream_t st) {
  FH_CHECK(S.kind == K_CSR, "csr_spmm: not a CSR matrix");
  const bool tr = (op != 'N');
  if (tr) FH_CHECK(S.has_csc, "csr_spmm: CSC twin missing");
  const int rows = tr ? S.n : S.m;
  if (rows == 0 || ncols == 0) return;
  const int* rp = tr ? S.cp.as<int>() : S.rp.as<int>();
  const int* ci = tr ? S.ri.as<int>() : S.ci.as<int>();
  const double* v = tr ? S.cv.as<double>() : S.v.as<double>();
  dim3 grid((rows + 63) / 64, (ncols + 3) / 4);
  if (transX)
    hipLaunchKernelGGL(k_csr_spmm_f64<true>, grid, dim3(256), 0, st, rows, ncols, rp, ci, v, X, ldx, Y, ldy, alpha,
                       beta);
  else
    hipLaunchKernelGGL(k_csr_spmm_f64<false>, grid, dim3(256), 0, st, rows, ncols, rp, ci, v, X, ldx, Y, ldy, alpha,
                       beta);
  FH_LAUNCH_CHECK();
}

void csr_spmm_f64(const DMat& S, char op, const double* X, int ldx, int ncols, double* Y, int ldy, double alpha,
                  double beta, hipStream_t st) {
  csr_spmm_f64_x(S, op, X, ldx, 0, ncols, Y, ldy, alpha, beta, st);
}

// C(i,j) = alpha * sum_{k in col j of op(S)} op(A)(i,k) * op(S)(k,j) + beta*C(i,j)
// (cp, ri, cv) is the column-compressed form of op(S): CSC of S for op N,
// CSR of S for op T.
template <bool TRANSA>
__global__ __launch_bounds__(256) void k_dense_csr_f64(int m, int n, const double* __restrict__ A, int lda,
                                                        const int* __restrict__ cp,
                                                        const int* __restrict__ ri, const double* __restrict__ cv,
                                                        double* __restrict__ C, int ldc, double alpha, double beta) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= m || j >= n) return;
  const int b = cp[j], e = cp[j + 1];
  double acc = 0.0;
  const double* ai = A + (size_t)i * lda;
  for (int base = b; base < e; base += 8) {
    const int nb = min(8, e - base);
    int rr[8];
    double ww[8], aa[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int p = base + min(q, nb - 1);
      rr[q] = ri[p];
      ww[q] = cv[p];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) aa[q] = TRANSA ? ai[rr[q]] : A[(size_t)rr[q] * lda + i];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double f = fma(aa[q], ww[q], acc);
      acc = q < nb ? f : acc;
    }
  }
  const double t = alpha * acc;
  double* c = C + (size_t)j * ldc + i;
  *c = (beta == 0.0) ? t : fma(beta, *c, t);
}

void dense_csr_f64(const double* A, int lda, char opA, int m, const DMat& S, char opS, double* C, int ldc,
                   double alpha, double beta, hipStream_t st) {
  FH_CHECK(S.kind == K_CSR, "dense_csr: not a CSR matrix");
  const bool trS = (opS != 'N');
  if (!trS) FH_CHECK(S.has_csc, "dense_csr: CSC twin missing");
  const int n = trS ? S.m : S.n;  // columns of op(S)
  if (m == 0 || n == 0) return;
  const int* cp = trS ? S.rp.as<int>() : S.cp.as<int>();
  const int* ri = trS ? S.ci.as<int>() : S.ri.as<int>();
  const double* cv = trS ? S.v.as<double>() : S.cv.as<double>();
  dim3 grid((m + 63) / 64, (n + 3) / 4);
  if (opA != 'N')
    hipLaunchKernelGGL(k_dense_csr_f64<true>, grid, dim3(256), 0, st, m, n, A, lda, cp, ri, cv, C, ldc, alpha, beta);
  else
    hipLaunchKernelGGL(k_dense_csr_f64<false>, grid, dim3(256), 0, st, m, n, A, lda, cp, ri, cv, C, ldc, alpha, beta);
  FH_LAUNCH_CHECK();
}

}  // namespace fh
