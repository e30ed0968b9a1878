// fh_internal.hpp -- shared host/device declarations of libfaust_hip.
//
// Numerics of the FP64 path ("canonical order", shared with the oracle):
//   * a product entry is acc = 0; for k ascending: acc = fma(a_ik, b_kj, acc);
//     terms with an exact-zero operand may be skipped (fma(0,b,acc) == acc for
//     finite b), so sparse and dense formulations give the same bits;
//   * epilogue t = alpha*acc; out = (beta == 0) ? t : fma(beta, c_old, t);
//   * sums/dots: 4096-element chunks, 256 lanes per chunk (lane t takes
//     elements t, t+256, ...), a halving tree p[t] = p[t] + p[t+s] for
//     s = 128..1, then chunk partials added in chunk order by one thread.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/faust_hip.h"

namespace fh {

constexpr int RED_CHUNK = 4096;
constexpr int RED_LANES = 256;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& s);

#define FH_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw ::fh::Error(FH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

#define FH_CHECK(cond, msg)                                \
  do {                                                     \
    if (!(cond)) throw ::fh::Error(FH_ERR_ARG, (msg));     \
  } while (0)

#define FH_LAUNCH_CHECK() FH_HIP(hipGetLastError())

// ------------------------------------------------------------ device memory
// A size-class caching pool: freed blocks are kept for reuse so that the PALM
// and chain loops do not call hipMalloc/hipFree per product.
void* dev_alloc(size_t bytes);
void dev_free(void* p);
void pool_trim();

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  explicit DBuf(size_t b) { alloc(b); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t b) {
    release();
    if (b) p = dev_alloc(b);
    bytes = b;
  }
  void release() {
    if (p) dev_free(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

inline size_t dtype_size(int dt) { return dt == FH_F64 ? 8 : dt == FH_F32 ? 4 : 2; }

// ------------------------------------------------------------ matrices
enum Kind { K_DENSE = 0, K_CSR = 1, K_BSR = 2 };

// A device matrix.  Dense: column-major, ld = m.  CSR: int32 rowptr/colind,
// columns sorted inside each row, with its CSC twin (rows ascending inside
// each column) built when the matrix is created; op T/H products walk the
// twin.  BSR: see below and k_mixed.hip.
struct DMat {
  Kind kind = K_DENSE;
  int dtype = FH_F64;
  int m = 0, n = 0;
  DBuf d;                 // dense values
  DBuf rp, ci, v;         // CSR
  int64_t nnz = 0;        // CSR: stored entries; dense: -1 until counted
  int max_row = -1, max_col = -1;  // CSR uploaded from the host: widest row / column (-1: unknown)
  mutable bool has_csc = false;
  mutable DBuf cp, ri, cv; // CSC
  // BSR (32x32 blocks): bdat holds each block in MFMA A-operand order, entry
  // (i, k) at ((k/16)*64 + ((k/8)%2)*32 + i)*8 + k%8, in BF16; FP32 values
  // are stored as a BF16 pair per entry (hi plane then lo plane per block,
  // hi = bf16(v), lo = bf16(v - hi)).  t* is the BSR of the transpose.
  int bm = 0, bn = 0, bnnz = 0;
  bool split = false;
  DBuf brp, bci, bdat;
  DBuf tbrp, tbci, tbdat;
  // ELL view of the block rows (and of the transpose): ell_w slots per block
  // row (a multiple of 4), slot = (block index into bdat, block column); an
  // empty slot points at the all-zero block bdat holds after its bnnz blocks.
  // ell_w == 0 when the widest block row is too wide for it.
  int ell_w = 0, tell_w = 0;
  DBuf ell, tell;
  // FP32 handles also keep their blocks in FP32 (k_bsr_f32's operand order:
  // lane l = 32*h + r holds row r, columns 16*h .. 16*h + 15 of the block, 16
  // consecutive floats), in bdat's (and tbdat's) block order, zero block last
  DBuf f32dat, tf32dat;
  // the blocks again in ELL order: slot (br, q) at (br * ell_w + q) blocks
  // (zero blocks in empty slots), so a kernel finds a slot's block without
  // loading the slot first; tebdat likewise for the transpose
  DBuf ebdat, tebdat;
  // CSR uploaded from the host (FP32 / FP64): ELL view of the rows (and of
  // the transpose's rows) for the gather kernels, csr_ell_w slots per row
  // (the widest row rounded up to 4 for FP32, at most 16; to 8 for FP64, at
  // most 32; 0: none), slot q of row r at r * w + q, unused slots with
  // column -1 and value 0
  int csr_ell_w = 0, csr_tell_w = 0;
  DBuf csr_ell_ci, csr_ell_v, csr_tell_ci, csr_tell_v;
  bool is_identity = false;
  // CSR copy of a dense matrix that turned out to be mostly zeros (built by
  // the product dispatcher; same values, so products keep the same bits)
  mutable std::shared_ptr<DMat> csr_twin;
  mutable bool sparsity_checked = false;
};
using DMatP = std::shared_ptr<DMat>;

// ------------------------------------------------------------ product modes
// Multiplication mode of the Faust product running on this thread (set from
// the Faust handle by fh_faust_multiply, TransformHelper::set_FM_mul_mode's
// role): FH_MUL_AUTO lets the defaults and the FH_* environment knobs decide,
// the other FH_MUL_* values of faust_hip.h force one execution form.
int mul_mode();
struct MulModeScope {
  int prev;
  explicit MulModeScope(int m);
  ~MulModeScope();
};

// ------------------------------------------------------------ kernels (FP64 canonical)
// Y = alpha*op(S)*X + beta*Y for CSR S, dense col-major X/Y.
void csr_spmm_f64(const DMat& S, char op, const double* X, int ldx, int ncols, double* Y, int ldy, double alpha,
                  double beta, hipStream_t st);
// Whole FP64 CSR chain in one launch when every intermediate column panel
// fits a workgroup's LDS: Y = alpha * op(S_1) ... op(S_J) X + beta*Y, factors
// given in application order (the first one applied to X comes first), each
// op'd factor as (row pointer, column index, value, rows, nnz, whether the
// factor-by-factor kernel reads it through an ELL view).  plan returns the
// columns per workgroup, or 0 when the shape does not fit or the
// factor-by-factor path is the faster (then use that path).
struct ChainCsrOp {
  const int* rp;
  const int* ci;
  const double* v;
  int rows;
  int nnz;
  int ell;
};
int csr_chain_plan_f64(const std::vector<ChainCsrOp>& ops, int in_rows, int p);
void csr_chain_run_f64(const std::vector<ChainCsrOp>& ops, int in_rows, int wc, const double* X, int ldx, int p,
                       double* Y, int ldy, double alpha, double beta, hipStream_t st);
// The same chain in one launch with every row of every factor and of X at
// most 1024 long (one thread per row, the first pf = 8 or 16 entries of a
// row loaded two factors ahead): fits says whether the shape allows it,
// plan returns the default columns per workgroup or 0 (mode / environment),
// default_wc the default columns per workgroup for p columns.
bool csr_chain_pf_fits(const std::vector<ChainCsrOp>& ops, int in_rows, int p);
int csr_chain_pf_plan_f64(const std::vector<ChainCsrOp>& ops, int in_rows, int p);
int csr_chain_pf_default_wc(int p);
void csr_chain_pf_run_f64(const std::vector<ChainCsrOp>& ops, int in_rows, int wc, int pf, const double* X, int ldx,
                          int p, double* Y, int ldy, double alpha, double beta, hipStream_t st);
// Columns per workgroup of the staged one-launch chain for this shape
// (whatever the chain length), or 0 when it does not fit.
int csr_chain_fit_wc(const std::vector<ChainCsrOp>& ops, int in_rows, int p);
// C = alpha*op(A)*op(S) + beta*C, A dense col-major, S CSR.
void dense_csr_f64(const double* A, int lda, char opA, int m, const DMat& S, char opS, double* C, int ldc,
                   double alpha, double beta, hipStream_t st);
// C = alpha*op(A)*op(B) + beta*C, all dense col-major.
void gemm_f64(char opA, char opB, int m, int n, int k, double alpha, const double* A, int lda, const double* B,
              int ldb, double beta, double* C, int ldc, hipStream_t st);

// canonical reductions; result written to a device double
void reduce_dot_f64(const double* x, const double* y, int64_t n, double* out, hipStream_t st);
void reduce_sum_f64(const double* x, int64_t n, double* out, hipStream_t st);
void diag_dot_f64(const double* A, const double* S, int m, int n, double* diag, hipStream_t st);
// dense -> CSR (drops exact zeros), nnz returned on host (synchronizes)
void dense_to_csr_f64(const double* D, int m, int n, DMat& out, hipStream_t st);
void csr_to_dense_f64(const DMat& S, double* D, hipStream_t st);
int64_t count_nnz_f64(const double* D, int64_t n, hipStream_t st);
// nonzero entries of n values of type dtype (FH_F64, FH_F32 or FH_BF16 bits)
int64_t count_nnz(const void* D, int dtype, int64_t n, hipStream_t st);
// elementwise helpers
void axpby_f64(double* y, const double* x, int64_t n, double a, double b, hipStream_t st); // y = a*x + b*y canonical
void scale_f64(double* x, int64_t n, const double* dscale, double hscale, int mode, hipStream_t st);
void sub_f64(double* y, const double* x, int64_t n, hipStream_t st);  // y = y - x
void fill_f64(double* x, int64_t n, double v, hipStream_t st);
void eye_f64(double* x, int m, int n, hipStream_t st);
// projectors on dense col-major M (k per row/col/global); exact partial_sort tie semantics
void prox_f64(int name, double* M, int m, int n, int64_t k, bool normalized, bool pos, const double* supp,
              hipStream_t st);
// normcol / normlin: every column / row scaled to 2-norm s (faust_prox.hpp:205-255)
void prox_real_f64(int name, double* M, int m, int n, double s, bool normalized, bool pos, hipStream_t st);

// power iteration over a chain of (factor, op) applied last to first.
struct ChainOp {
  const DMat* M;
  char op;
};
double power_iteration_f64(const std::vector<ChainOp>& ops, int n, int maxit, double thr, int* flag,
                           hipStream_t st);
// two independent chains at once (one workgroup each when both fit the
// persistent kernel, else one after the other); the same bits as two calls of
// power_iteration_f64; flag receives the second chain's
void power_iteration_pair_f64(const std::vector<ChainOp>& ops1, int n1, const std::vector<ChainOp>& ops2, int n2,
                              int maxit, double thr, double* lam1, double* lam2, int* flag, hipStream_t st);

// host-side start vector: srand(0xF4+'u'+57), Random on [-1, 1]
const std::vector<double>& power_x0(int n);

// generic FP64 product dispatch: out = alpha*op(A)*op(B) + beta*out (dense out)
void matmul_f64(const DMat& A, char opA, const DMat& B, char opB, double alpha, double beta, DMat& out,
                hipStream_t st);
DMatP dense_of(const DMat& A, hipStream_t st);
int64_t nnz_of(DMat& A, hipStream_t st);

// ------------------------------------------------------------ butterfly
// One split of a butterfly node (k_butterfly.hip): node [a, b) of L supports,
// n x 2^in_logB values (row r, entry Q at column
// (r & ~(2^(in_logB+in_logs) - 1)) + (r & (2^in_logs - 1)) + 2^in_logs Q; the
// root [0, L) is the whole matrix row-major); X receives the node [a, c)
// (logBx = c - a, logsx = L - c), Y the node [c, b).
struct BflyStep {
  const double* in = nullptr;
  int in_logB = 0, in_logs = 0;
  int n = 0;
  int logBx = 0, logsx = 0, logBy = 0, logsy = 0;
  double* X = nullptr;
  double* Y = nullptr;
};
void bfly_split(const BflyStep& s, hipStream_t st);
bool bfly_block_fits(int logBx, int logBy);
void bfly_scatter_dense(const double* v, int n, int logs, double* out, hipStream_t st);
// CSR leaf factor (CSC twin and ELL views included) from a 2-entry-per-row
// node on the device; colmap / colinv (device, both or neither): relabel
// column c as colmap[c]
void bfly_leaf_csr(const double* v, int n, int logs, const int* colmap, const int* colinv, DMat& out,
                   hipStream_t st);
// root node (row-major n x n) = A with columns permuted, root(r, c) = A(r, pinv[c]); A column-major (ld lda)
// or row-major (row_major, ld lda), device memory; pinv_dev device or nullptr
void bfly_root(const double* A, size_t lda, int row_major, int n, const int* pinv_dev, double* out, hipStream_t st);

}  // namespace fh
