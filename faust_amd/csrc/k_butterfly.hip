// k_butterfly.hip -- the lifting step of the butterfly factorization.
//
// A butterfly node factor covering supports [a, b) of L has the support
// kron(I_{2^a}, kron(ones(B, B), I_s)), B = 2^(b-a), s = 2^(L-b): row r holds
// exactly B entries, entry Q at column (r & ~(B s - 1)) + (r & (s - 1)) + s Q.
// It is kept as n x B values, row-major (value (r, Q) at r * B + Q).
//
// Splitting node [a, b) at c (lifting_two_layers_factorization,
// faust_butterfly.hpp:176-246) factors it as X Y with X on [a, c) and Y on
// [c, b).  For each t in [0, n) the rows t touches in X's support and the
// columns t touches in Y's support select a Bx x By block M_t of the node
// (Bx = 2^(c-a), By = 2^(b-c)); with M_t ~ sigma u v^T its best rank-1
// approximation, column t of X gets u sqrt(sigma) and row t of Y gets
// sqrt(sigma) v^T (MatDense::best_low_rank with rank 1).  The blocks of
// different t are disjoint, so every block is one independent task:
//   * blocks with a side of 2 (every step of the 'left' / 'right' drivers and
//     the bottom of the tree): closed form from the 2 x 2 Gram matrix, one
//     wave per block, any length;
//   * larger blocks: one-sided (Hestenes) Jacobi on the block held in LDS,
//     one workgroup per block, rounds of disjoint column pairs in
//     round-robin order, until a sweep rotates nothing.
// All arithmetic is FP64.  Singular vectors are defined up to sign; the sign
// taken here is whatever the rotations leave.
#include <cfloat>

#include "fh_internal.hpp"

namespace fh {

constexpr int BF_THREADS = 256;
constexpr int BF_MAX_SWEEPS = 30;

__device__ __forceinline__ double bf_in(const BflyStep& s, int r, int c) {
  if (s.in_dense) return s.in[(size_t)r + (size_t)c * s.n];
  const int bs = 1 << (s.in_logB + s.in_logs);
  return s.in[((size_t)r << s.in_logB) + ((c & (bs - 1)) >> s.in_logs)];
}

// row of X's support that block t's q-th row is; column of Y's support that
// its q-th column is
__device__ __forceinline__ int bf_row(const BflyStep& s, int t, int q) {
  const int bs = 1 << (s.logBx + s.logsx);
  return (t & ~(bs - 1)) + (t & ((1 << s.logsx) - 1)) + (q << s.logsx);
}
__device__ __forceinline__ int bf_col(const BflyStep& s, int t, int q) {
  const int bs = 1 << (s.logBy + s.logsy);
  return (t & ~(bs - 1)) + (t & ((1 << s.logsy) - 1)) + (q << s.logsy);
}

// W0(j, i): the block seen as d columns of length D.  tr (Bx < By): the
// columns are the Bx rows of M_t; otherwise the By columns of M_t.
__device__ __forceinline__ double bf_w0(const BflyStep& s, bool tr, int t, int j, int i) {
  return tr ? bf_in(s, bf_row(s, t, j), bf_col(s, t, i)) : bf_in(s, bf_row(s, t, i), bf_col(s, t, j));
}

// write the rank-1 pair: along-length values (D) and along-column values (d)
__device__ __forceinline__ void bf_put_len(const BflyStep& s, bool tr, int t, int i, double v) {
  if (tr) {
    s.Y[((size_t)t << s.logBy) + i] = v;
  } else {
    const int bs = 1 << (s.logBx + s.logsx);
    s.X[((size_t)bf_row(s, t, i) << s.logBx) + ((t & (bs - 1)) >> s.logsx)] = v;
  }
}
__device__ __forceinline__ void bf_put_col(const BflyStep& s, bool tr, int t, int j, double v) {
  if (tr) {
    const int bs = 1 << (s.logBx + s.logsx);
    s.X[((size_t)bf_row(s, t, j) << s.logBx) + ((t & (bs - 1)) >> s.logsx)] = v;
  } else {
    s.Y[((size_t)t << s.logBy) + j] = v;
  }
}

// Jacobi rotation that makes columns p, q orthogonal, from a = |w_p|^2,
// b = |w_q|^2, g = w_p . w_q (g != 0): w_p' = c w_p - sn w_q,
// w_q' = sn w_p + c w_q, |w_p'|^2 = a - tn g, |w_q'|^2 = b + tn g.
__device__ __forceinline__ void bf_rot(double a, double b, double g, double& c, double& sn, double& tn) {
  const double zeta = (b - a) / (2.0 * g);
  const double az = fabs(zeta);
  // zeta^2 overflows past 1e154: tn -> 1 / (2 zeta) there
  const double root = az < 1e150 ? sqrt(1.0 + az * az) : az;
  tn = (zeta >= 0.0 ? 1.0 : -1.0) / (az + root);
  c = 1.0 / sqrt(1.0 + tn * tn);
  sn = c * tn;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------- d = 2
__global__ __launch_bounds__(BF_THREADS) void k_bfly_d2(BflyStep s) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int t = blockIdx.x * (BF_THREADS / 64) + w;
  if (t >= s.n) return;  // whole wave
  const bool tr = s.logBx < s.logBy;
  const int D = 1 << (tr ? s.logBy : s.logBx);
  double a = 0.0, b = 0.0, g = 0.0;
  for (int i = lane; i < D; i += 64) {
    const double w0 = bf_w0(s, tr, t, 0, i), w1 = bf_w0(s, tr, t, 1, i);
    a = fma(w0, w0, a);
    b = fma(w1, w1, b);
    g = fma(w0, w1, g);
  }
  a = wave_sum(a);
  b = wave_sum(b);
  g = wave_sum(g);
  // top eigenpair (lam, (e0, e1)) of [[a, g], [g, b]]
  double e0, e1, lam;
  if (g == 0.0) {
    if (a >= b) { e0 = 1.0; e1 = 0.0; lam = a; }
    else { e0 = 0.0; e1 = 1.0; lam = b; }
  } else {
    double c, sn, tn;
    bf_rot(a, b, g, c, sn, tn);
    const double ap = a - tn * g, bp = b + tn * g;
    if (bp >= ap) { e0 = sn; e1 = c; lam = bp; }
    else { e0 = c; e1 = -sn; lam = ap; }
  }
  double kl = 0.0, kc = 0.0;
  if (lam > 0.0) {
    const double sigma = sqrt(lam), rs = sqrt(sigma);
    kl = 1.0 / rs;
    kc = 1.0 / (sigma * rs);
  }
  // w* = e0 w0 + e1 w1 (= sigma times the unit singular vector on the long
  // side); the short side is W0^T w* = G e
  for (int i = lane; i < D; i += 64) {
    const double w0 = bf_w0(s, tr, t, 0, i), w1 = bf_w0(s, tr, t, 1, i);
    bf_put_len(s, tr, t, i, (e0 * w0 + e1 * w1) * kl);
  }
  if (lane == 0) bf_put_col(s, tr, t, 0, (e0 * a + e1 * g) * kc);
  if (lane == 1) bf_put_col(s, tr, t, 1, (e0 * g + e1 * b) * kc);
}

// ---------------------------------------------------------------- d >= 4
// LDS: d columns of D doubles, column stride D + 1.
__global__ __launch_bounds__(BF_THREADS) void k_bfly_jacobi(BflyStep s) {
  extern __shared__ double W[];
  __shared__ double nrm[128];
  __shared__ int flag, top;
  const int t = blockIdx.x, tid = threadIdx.x;
  const bool tr = s.logBx < s.logBy;
  const int logD = tr ? s.logBy : s.logBx, logd = tr ? s.logBx : s.logBy;
  const int D = 1 << logD, d = 1 << logd, ld = D + 1;
  for (int e = tid; e < (D << logd); e += BF_THREADS) {
    const int j = e >> logD, i = e & (D - 1);
    W[j * ld + i] = bf_w0(s, tr, t, j, i);
  }
  if (tid == 0) flag = 0;
  __syncthreads();
  const int npair = d >> 1;
  const int g = min(64, BF_THREADS / npair);  // lanes per pair, a power of two
  const int k = tid / g, lg = tid % g;
  const bool active = k < npair;
  const double tol = (double)D * DBL_EPSILON;
  for (int sweep = 0; sweep < BF_MAX_SWEEPS; ++sweep) {
    for (int r = 0; r < d - 1; ++r) {
      if (active) {
        // round r of the circle schedule: player d-1 meets r, the others
        // meet their mirror image about r (mod d-1)
        int p, q;
        if (k == 0) {
          p = r;
          q = d - 1;
        } else {
          const int m = d - 1;
          const int x = (r + k) % m, y = (r - k + m) % m;
          p = min(x, y);
          q = max(x, y);
        }
        double* wp = W + p * ld;
        double* wq = W + q * ld;
        double a = 0.0, b = 0.0, gm = 0.0;
        for (int i = lg; i < D; i += g) {
          const double x0 = wp[i], x1 = wq[i];
          a = fma(x0, x0, a);
          b = fma(x1, x1, b);
          gm = fma(x0, x1, gm);
        }
        for (int o = g >> 1; o >= 1; o >>= 1) {
          a += __shfl_xor(a, o, 64);
          b += __shfl_xor(b, o, 64);
          gm += __shfl_xor(gm, o, 64);
        }
        if (fabs(gm) > tol * sqrt(a * b)) {
          double c, sn, tn;
          bf_rot(a, b, gm, c, sn, tn);
          for (int i = lg; i < D; i += g) {
            const double x0 = wp[i], x1 = wq[i];
            wp[i] = c * x0 - sn * x1;
            wq[i] = sn * x0 + c * x1;
          }
          if (lg == 0) flag = 1;
        }
      }
      __syncthreads();
    }
    const int any = flag;
    __syncthreads();
    if (tid == 0) flag = 0;
    __syncthreads();
    if (!any) break;
  }
  // column norms; the largest is sigma^2
  if (active) {
    for (int h = 0; h < 2; ++h) {
      const int j = k + h * npair;
      double a = 0.0;
      for (int i = lg; i < D; i += g) a = fma(W[j * ld + i], W[j * ld + i], a);
      for (int o = g >> 1; o >= 1; o >>= 1) a += __shfl_xor(a, o, 64);
      if (lg == 0) nrm[j] = a;
    }
  }
  __syncthreads();
  if (tid == 0) {
    int b = 0;
    for (int j = 1; j < d; ++j)
      if (nrm[j] > nrm[b]) b = j;
    top = b;
  }
  __syncthreads();
  const double lam = nrm[top];
  double kl = 0.0, kc = 0.0;
  if (lam > 0.0) {
    const double sigma = sqrt(lam), rs = sqrt(sigma);
    kl = 1.0 / rs;
    kc = 1.0 / (sigma * rs);
  }
  const double* ws = W + top * ld;
  for (int i = tid; i < D; i += BF_THREADS) bf_put_len(s, tr, t, i, ws[i] * kl);
  // short side: W0^T w*, from the block as it was
  if (active) {
    for (int h = 0; h < 2; ++h) {
      const int j = k + h * npair;
      double a = 0.0;
      for (int i = lg; i < D; i += g) a = fma(bf_w0(s, tr, t, j, i), ws[i], a);
      for (int o = g >> 1; o >= 1; o >>= 1) a += __shfl_xor(a, o, 64);
      if (lg == 0) bf_put_col(s, tr, t, j, a * kc);
    }
  }
}

// dense n x n column-major factor from a 2-entry-per-row node (B = 2, s = 2^logs)
__global__ void k_bfly_scatter(const double* v, int n, int logs, double* out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int bs = 2 << logs;
  const int c0 = (r & ~(bs - 1)) + (r & ((1 << logs) - 1));
  out[(size_t)r + (size_t)c0 * n] = v[2 * (size_t)r];
  out[(size_t)r + (size_t)(c0 + (1 << logs)) * n] = v[2 * (size_t)r + 1];
}

static size_t bfly_lds_budget() {
  static size_t budget = 0;
  if (budget) return budget;
  int dev = 0, v = 0;
  size_t b = 64 * 1024;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > (int)b) {
    // the kernel's static LDS (norms and flags) comes on top of the dynamic part
    const int want = v - 2048;
    if (hipFuncSetAttribute((const void*)k_bfly_jacobi, hipFuncAttributeMaxDynamicSharedMemorySize, want) ==
        hipSuccess)
      b = (size_t)want;
  }
  (void)hipGetLastError();  // a refused attribute leaves the 64 KiB budget, not an error
  budget = b;
  return budget;
}

size_t bfly_block_lds_bytes(int logBx, int logBy) {
  const int logD = std::max(logBx, logBy), logd = std::min(logBx, logBy);
  return sizeof(double) * (size_t)((1 << logD) + 1) * (size_t)(1 << logd);
}

bool bfly_block_fits(int logBx, int logBy) {
  if (std::min(logBx, logBy) <= 1) return true;  // closed form, any length
  if (std::min(logBx, logBy) > 7) return false;  // column norms: 128 slots
  return bfly_block_lds_bytes(logBx, logBy) <= bfly_lds_budget();
}

void bfly_split(const BflyStep& s, hipStream_t st) {
  FH_CHECK(s.n >= 4 && (s.n & (s.n - 1)) == 0, "butterfly: order must be a power of two >= 4");
  FH_CHECK(s.logBx >= 1 && s.logBy >= 1, "butterfly: empty split");
  FH_CHECK(s.logBx + s.logsx == s.in_logB + s.in_logs && s.logsy == s.in_logs && s.logBy + s.logsy == s.logsx,
           "butterfly: split does not match its node");
  FH_CHECK(s.in_dense ? (s.in_logB + s.in_logs == (int)std::log2((double)s.n) && s.in_logs == 0) : true,
           "butterfly: dense input must be the whole matrix");
  if (std::min(s.logBx, s.logBy) == 1) {
    const int per = BF_THREADS / 64;
    hipLaunchKernelGGL(k_bfly_d2, dim3((s.n + per - 1) / per), dim3(BF_THREADS), 0, st, s);
  } else {
    FH_CHECK(bfly_block_fits(s.logBx, s.logBy), "butterfly: block too large for one workgroup's LDS");
    const size_t lds = bfly_block_lds_bytes(s.logBx, s.logBy);
    hipLaunchKernelGGL(k_bfly_jacobi, dim3(s.n), dim3(BF_THREADS), lds, st, s);
  }
  FH_LAUNCH_CHECK();
}

void bfly_scatter_dense(const double* v, int n, int logs, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_bfly_scatter, dim3((n + 255) / 256), dim3(256), 0, st, v, n, logs, out);
  FH_LAUNCH_CHECK();
}

}  // namespace fh
