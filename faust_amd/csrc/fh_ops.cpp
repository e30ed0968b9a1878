// fh_ops.cpp -- FP64 product dispatch over dense / CSR device matrices.
//
// Every path produces the canonical bits (one fma chain per entry, k
// ascending, zero terms skipped), so the dispatcher is free to pick the
// cheapest formulation for the operand formats: CSR x dense, dense x CSR, or
// dense GEMM.
#include <algorithm>

#include "fh_internal.hpp"

namespace fh {

void csr_spmm_f64_x(const DMat& S, char op, const double* X, int ldx, int transX, int ncols, double* Y, int ldy,
                    double alpha, double beta, hipStream_t st);

DMatP dense_of(const DMat& A, hipStream_t st) {
  auto D = std::make_shared<DMat>();
  D->kind = K_DENSE;
  D->dtype = FH_F64;
  D->m = A.m;
  D->n = A.n;
  D->d.alloc(sizeof(double) * std::max<size_t>((size_t)A.m * A.n, 1));
  if (A.kind == K_DENSE) {
    FH_HIP(hipMemcpyAsync(D->d.p, A.d.p, sizeof(double) * (size_t)A.m * A.n, hipMemcpyDeviceToDevice, st));
    D->nnz = A.nnz;
  } else if (A.kind == K_CSR) {
    csr_to_dense_f64(A, D->d.as<double>(), st);
    D->nnz = A.nnz;
  } else {
    throw Error(FH_ERR_UNSUPPORTED, "dense_of: BSR factors are not supported in the FP64 path");
  }
  return D;
}

int64_t nnz_of(DMat& A, hipStream_t st) {
  if (A.kind == K_CSR) return A.nnz;
  if (A.kind != K_DENSE) throw Error(FH_ERR_UNSUPPORTED, "nnz count: dense or CSR matrices only");
  if (A.nnz < 0) A.nnz = count_nnz(A.d.p, A.dtype, (int64_t)A.m * A.n, st);
  return A.nnz;
}

// A dense operand whose density is below 1/8 is multiplied through a CSR
// copy (the densified products of a few butterfly-like factors are such
// matrices).  Only matrices of at least 64K entries are examined.
static const DMat* sparse_view(const DMat& A, hipStream_t st) {
  if (A.kind != K_DENSE || A.dtype != FH_F64) return &A;
  const int64_t N = (int64_t)A.m * A.n;
  if (N < 65536) return &A;
  if (!A.sparsity_checked) {
    A.sparsity_checked = true;
    int64_t nz = nnz_of(const_cast<DMat&>(A), st);
    if (nz * 8 < N) {
      auto S = std::make_shared<DMat>();
      dense_to_csr_f64(A.d.as<double>(), A.m, A.n, *S, st);
      A.csr_twin = S;
    }
  }
  return A.csr_twin ? A.csr_twin.get() : &A;
}

// out = alpha*op(A)*op(B) + beta*out ; out is dense (allocated here when beta == 0)
void matmul_f64(const DMat& A0, char opA, const DMat& B0, char opB, double alpha, double beta, DMat& out,
                hipStream_t st) {
  const DMat& A = *sparse_view(A0, st);
  const DMat& B = (A.kind == K_CSR) ? B0 : *sparse_view(B0, st);
  const int M = (opA == 'N') ? A.m : A.n;
  const int K = (opA == 'N') ? A.n : A.m;
  const int KB = (opB == 'N') ? B.m : B.n;
  const int N = (opB == 'N') ? B.n : B.m;
  FH_CHECK(K == KB, "matmul: dimension mismatch");
  if (beta != 0.0) {
    FH_CHECK(out.kind == K_DENSE && out.m == M && out.n == N && out.d.p, "matmul: beta needs a dense output");
  } else {
    DBuf nd(sizeof(double) * std::max<size_t>((size_t)M * N, 1));
    out.d = std::move(nd);
    out.kind = K_DENSE;
    out.m = M;
    out.n = N;
  }
  out.dtype = FH_F64;
  out.nnz = -1;
  out.csr_twin.reset();
  out.sparsity_checked = false;
  double* C = out.d.as<double>();
  if (M == 0 || N == 0) return;
  if (A.kind == K_CSR) {
    // B must be read densely; op(B) = B^T is read transposed by the kernel
    DMatP Bd;
    const DMat* Bp = &B;
    if (B.kind != K_DENSE) {
      Bd = dense_of(B, st);
      Bp = Bd.get();
    }
    csr_spmm_f64_x(A, opA, Bp->d.as<double>(), Bp->m, opB != 'N' ? 1 : 0, N, C, M, alpha, beta, st);
  } else if (B.kind == K_CSR) {
    FH_CHECK(A.kind == K_DENSE, "matmul: unsupported operand format");
    dense_csr_f64(A.d.as<double>(), A.m, opA, M, B, opB, C, M, alpha, beta, st);
  } else {
    FH_CHECK(A.kind == K_DENSE && B.kind == K_DENSE, "matmul: unsupported operand format");
    gemm_f64(opA, opB, M, N, K, alpha, A.d.as<double>(), A.m, B.d.as<double>(), B.m, beta, C, M, st);
  }
}

}  // namespace fh
