// fh_batched.hpp -- job descriptors shared by the batched PALM4MSA kernels
// (k_batched.hip) and their host driver (fh_batched.cpp).
#pragma once
#include "fh_internal.hpp"

namespace fh {

constexpr int BATCH_NMAX = 2048;  // largest target edge (power-iteration vectors live in LDS)
constexpr int BGEMM_MFMA = 65;    // tile code of the FP32 MFMA grouped GEMM (64 x 64 tiles)

// C_out = alpha_b * op(A) op(B) + beta * C_in on n x n column-major blocks;
// alpha mode 0 -> 1, 1 -> lambda_b, 2 -> -lambda_b / c_b with
// c_b = 1.001 lambda_b^2 nR_b^2 nL_b^2; beta mode 0 -> 0, 1 -> -1, 2 -> +1.
template <class T>
struct BGemmJob {
  const T* A;
  const T* B;
  const T* Cin;
  T* Cout;
  int m, n, k;
  int transA, transB;
  int alpha_mode, beta_mode;
  int target;
};

struct BScal {
  double* lambda;
  double* nR;
  double* nL;
};

// sqrt(lambda_max(P P^T)) of an n x n block -> out[target]; P == nullptr
// writes 1 (no factor on that side).  vec: this slot's start vector.
template <class T>
struct BPowJob {
  const T* P;
  int n;
  int target;
  double* out;
  double* vec;
};

// A factor of one target as compressed rows and compressed columns (entries
// of a row in column order, of a column in row order), rebuilt from the
// dense factor after each projection; cap entries of room each.
template <class T>
struct BSpFac {
  int* rp;
  int* ci;
  T* v;
  int* cp;
  int* ri;
  T* cv;
  int cap;
  // the kept entries (column-major positions, in no order) the last sp
  // projection of this factor listed, and their count, or -1 when it listed
  // none (ties ranked by index, more than the list holds, or no projection yet)
  int* klist;
  int* kcount;
};

// Build job: the dense n x n factor M -> F.
template <class T>
struct BSpBuildJob {
  const T* M;
  BSpFac<T> F;
  int n;
};

// sqrt(lambda_max(P P^T)) with P = F_0 F_1 ... F_{m-1} given by its sparse
// factors (fac[0..m-1]); m == 0 writes 1.  vec: this slot's start vector.
template <class T>
struct BPowSpJob {
  const BSpFac<T>* fac;
  int m;
  int n;
  int target;
  double* out;
  double* vec;
};

// Cout = alpha_b * (op(S) X  or  X op(S)) + beta * Cin on n x n column-major
// blocks, S a sparse factor (alpha / beta modes as BGemmJob).  side 0: S on
// the left, output entry (i, j) from row i of op(S) gathering column j of X;
// side 1: S on the right, column j of the output from column j of op(S)
// combining columns of X.
template <class T>
struct BSpmmJob {
  BSpFac<T> S;
  const T* X;
  const T* Cin;
  T* Cout;
  int n;
  int side, transS;
  int alpha_mode, beta_mode;
  int target;
};

template <class T>
struct BProxJob {
  T* M;
  const T* mask;  // SUPP only
  int64_t len;
  int64_t k;
  int* klist;   // optional: where the kept entries are listed (room for kcap)
  int* kcount;  // their count, or -1
  int kcap;
};

template <class T>
struct BLamJob {
  const T* A;
  const T* S;
  int64_t len;
  int target;
};

template <class T>
void launch_bgemm(const BGemmJob<T>* jobs, const int3* tiles, int ntiles, int tile, const BScal& sc, hipStream_t st);
template <class T>
void launch_bpower(const BPowJob<T>* jobs, int njobs, int maxit, double thr, int warm, hipStream_t st);
template <class T>
void launch_bprox(const BProxJob<T>* jobs, int njobs, int mode, hipStream_t st);
template <class T>
void launch_bsparse(const BSpBuildJob<T>* jobs, int njobs, hipStream_t st);
// tiles: (job, first column, columns) per workgroup
template <class T>
void launch_bspmm(const BSpmmJob<T>* jobs, const int3* tiles, int ntiles, const BScal& sc, hipStream_t st);
template <class T>
// mc: 1 or 2 when every job has mc factors or none and nmax <= 512 (their rows
// are then held in registers), else 0
void launch_bpower_sp(const BPowSpJob<T>* jobs, int njobs, int nmax, int mc, int maxit, double thr, int warm,
                      hipStream_t st);
template <class T>
void launch_blambda(const BLamJob<T>* jobs, int njobs, double* lambda, hipStream_t st);
template <class T>
void launch_brelerr(const BLamJob<T>* jobs, int njobs, const double* lambda, double* out, hipStream_t st);
template <class T>
void launch_bcopy(T* dst, const T* src, int64_t n, hipStream_t st);
void launch_cast_f32_f64(double* dst, const float* src, int64_t n, hipStream_t st);
// dst_j = (double) src_j for every job j in one launch (src FP32 or FP64)
struct BCastJob {
  double* dst;
  const void* src;
  int64_t len;
};
void launch_bcast(const BCastJob* jobs, int njobs, bool src_f32, hipStream_t st);
void launch_mask_nonzero_f32(double* mask, const float* src, int64_t n, hipStream_t st);

}  // namespace fh
