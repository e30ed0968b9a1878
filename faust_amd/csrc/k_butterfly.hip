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
//   * larger blocks (both sides >= 4): the top eigenvector of the block's
//     Gram matrix by repeated squaring in LDS, one workgroup per block.
// All arithmetic is FP64.  Singular vectors are defined up to sign; the sign
// taken here is whatever the computation leaves.
#include <cfloat>
#include <cstdlib>

#include "fh_internal.hpp"

namespace fh {

constexpr int BF_THREADS = 256;

__device__ __forceinline__ double bf_in(const BflyStep& s, int r, int c) {
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

// top eigenpair of the 2 x 2 Gram matrix [[a, g], [g, b]] as the scales of
// the rank-1 pair: the length side gets (e0 w0 + e1 w1) kl, the column side
// (G e) kc
__device__ __forceinline__ void bf_d2_pair(double a, double b, double g, double& e0, double& e1, double& kl,
                                           double& kc) {
  double lam;
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
  kl = 0.0;
  kc = 0.0;
  if (lam > 0.0) {
    const double sigma = sqrt(lam), rs = sqrt(sigma);
    kl = 1.0 / rs;
    kc = 1.0 / (sigma * rs);
  }
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
  double e0, e1, kl, kc;
  bf_d2_pair(a, b, g, e0, e1, kl, kc);
  // w* = e0 w0 + e1 w1 (= sigma times the unit singular vector on the long
  // side); the short side is W0^T w* = G e
  for (int i = lane; i < D; i += 64) {
    const double w0 = bf_w0(s, tr, t, 0, i), w1 = bf_w0(s, tr, t, 1, i);
    bf_put_len(s, tr, t, i, (e0 * w0 + e1 * w1) * kl);
  }
  if (lane == 0) bf_put_col(s, tr, t, 0, (e0 * a + e1 * g) * kc);
  if (lane == 1) bf_put_col(s, tr, t, 1, (e0 * g + e1 * b) * kc);
}

// d = 2 with a long side of 2048..8192: one workgroup per block, which keeps
// its D x 2 values in registers (REG = D / 256 pairs per thread), so the
// block is read from HBM once.  The three sums are wave sums added in wave
// order, the same in every thread.
template <int REG>
__global__ __launch_bounds__(BF_THREADS) void k_bfly_d2_wide(BflyStep s) {
  __shared__ double red[3][BF_THREADS / 64];
  const int t = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool tr = s.logBx < s.logBy;
  double c0[REG], c1[REG];
  double a = 0.0, b = 0.0, g = 0.0;
#pragma unroll
  for (int k = 0; k < REG; ++k) {
    const int i = tid + k * BF_THREADS;
    c0[k] = bf_w0(s, tr, t, 0, i);
    c1[k] = bf_w0(s, tr, t, 1, i);
    a = fma(c0[k], c0[k], a);
    b = fma(c1[k], c1[k], b);
    g = fma(c0[k], c1[k], g);
  }
  a = wave_sum(a);
  b = wave_sum(b);
  g = wave_sum(g);
  if (lane == 0) {
    red[0][wave] = a;
    red[1][wave] = b;
    red[2][wave] = g;
  }
  __syncthreads();
  a = b = g = 0.0;
  for (int q = 0; q < BF_THREADS / 64; ++q) {
    a += red[0][q];
    b += red[1][q];
    g += red[2][q];
  }
  double e0, e1, kl, kc;
  bf_d2_pair(a, b, g, e0, e1, kl, kc);
#pragma unroll
  for (int k = 0; k < REG; ++k) bf_put_len(s, tr, t, tid + k * BF_THREADS, (e0 * c0[k] + e1 * c1[k]) * kl);
  if (tid == 0) bf_put_col(s, tr, t, 0, (e0 * a + e1 * g) * kc);
  if (tid == 1) bf_put_col(s, tr, t, 1, (e0 * g + e1 * b) * kc);
}

// ---------------------------------------------------------------- d >= 4
// Top singular pair of the block W0 (d columns of length D) from its Gram
// matrix G = W0^T W0 (d x d, in LDS): M <- G / tr G, then M <- M^2 / tr(M^2)
// until tr(M^2) reaches 1 (M is then the projector v v^T on the top
// eigenvector, the top right singular vector of W0), at most BF_MAX_SQ
// times.  v is M's column of largest diagonal entry, normalised; w = W0 v
// (= sigma u) comes from the block itself.  G and M carry relative errors of
// a few eps, which move v by about eps / (relative gap of sigma_1^2 and
// sigma_2^2).  Every thread sums the trace in the same order, so the
// stopping test is uniform.
constexpr int BF_MAX_SQ = 64;
constexpr int BF_CHUNK = 16;  // rows of W0 staged per Gram step (fewer when d < 16)

__device__ __forceinline__ int bf_block_index(int b, int nb) {
  // consecutive blocks share rows of the node: keep them on one XCD (L2);
  // workgroups are dealt to the 8 XCDs round-robin
  const int per = nb >> 3;
  return per > 0 && (nb & 7) == 0 ? (b & 7) * per + (b >> 3) : b;
}

// TS x TS register tile of a d x d product per thread (TS = 4 up to d = 64,
// 8 for d = 128; (d / TS)^2 <= 256 threads).  LDS: M (d x d), then ch x d
// staged rows of W0 for the Gram step.  A square is formed in registers and
// written over M after a barrier, so one d x d buffer suffices.
template <int TS>
__global__ __launch_bounds__(BF_THREADS) void k_bfly_gram(BflyStep s) {
  extern __shared__ double lds[];
  __shared__ double red[BF_THREADS / 64];
  const int t = bf_block_index(blockIdx.x, gridDim.x), tid = threadIdx.x;
  const bool tr = s.logBx < s.logBy;
  const int logD = tr ? s.logBy : s.logBx, logd = tr ? s.logBx : s.logBy;
  const int D = 1 << logD, d = 1 << logd;
  double* M = lds;          // d x d
  double* B = lds + d * d;  // ch x d staged rows of W0
  const int td = d / TS;    // tiles per side
  const bool tile = tid < td * td;
  const int p0 = (tid / td) * TS, q0 = (tid % td) * TS;
  double acc[TS][TS];
  // G = W0^T W0
#pragma unroll
  for (int a = 0; a < TS; ++a)
#pragma unroll
    for (int b = 0; b < TS; ++b) acc[a][b] = 0.0;
  const int ch = min(BF_CHUNK, D);
  for (int i0 = 0; i0 < D; i0 += ch) {
    for (int e = tid; e < ch * d; e += BF_THREADS) {
      const int r = e >> logd, j = e & (d - 1);
      B[r * d + j] = bf_w0(s, tr, t, j, i0 + r);
    }
    __syncthreads();
    if (tile) {
      for (int r = 0; r < ch; ++r) {
        double x[TS], y[TS];
#pragma unroll
        for (int a = 0; a < TS; ++a) {
          x[a] = B[r * d + p0 + a];
          y[a] = B[r * d + q0 + a];
        }
#pragma unroll
        for (int a = 0; a < TS; ++a)
#pragma unroll
          for (int b = 0; b < TS; ++b) acc[a][b] = fma(x[a], y[b], acc[a][b]);
      }
    }
    __syncthreads();
  }
  if (tile) {
#pragma unroll
    for (int a = 0; a < TS; ++a)
#pragma unroll
      for (int b = 0; b < TS; ++b) M[(p0 + a) * d + q0 + b] = acc[a][b];
  }
  __syncthreads();
  double tau = 0.0;
  for (int j = 0; j < d; ++j) tau += M[j * d + j];
  if (tau > 0.0) {
    double scale = 1.0 / tau;  // M holds the current matrix / scale
    for (int k = 0; k < BF_MAX_SQ; ++k) {
      const double sc2 = scale * scale;
      if (tile) {
#pragma unroll
        for (int a = 0; a < TS; ++a)
#pragma unroll
          for (int b = 0; b < TS; ++b) acc[a][b] = 0.0;
        for (int r = 0; r < d; ++r) {
          double x[TS], y[TS];
#pragma unroll
          for (int a = 0; a < TS; ++a) {
            x[a] = M[r * d + p0 + a];  // symmetric: row r = column r
            y[a] = M[r * d + q0 + a];
          }
#pragma unroll
          for (int a = 0; a < TS; ++a)
#pragma unroll
            for (int b = 0; b < TS; ++b) acc[a][b] = fma(x[a], y[b], acc[a][b]);
        }
      }
      __syncthreads();  // every tile of the square is formed before M is overwritten
      if (tile) {
#pragma unroll
        for (int a = 0; a < TS; ++a)
#pragma unroll
          for (int b = 0; b < TS; ++b) M[(p0 + a) * d + q0 + b] = acc[a][b] * sc2;
      }
      __syncthreads();
      double tr2 = 0.0;
      for (int j = 0; j < d; ++j) tr2 += M[j * d + j];
      scale = 1.0 / tr2;
      if (tr2 >= 1.0 - 16.0 * DBL_EPSILON) break;  // tr(M^2) = 1: M is rank one
    }
  }
  __syncthreads();
  // v: the column of largest diagonal entry, normalised
  int jb = 0;
  for (int j = 1; j < d; ++j)
    if (M[j * d + j] > M[jb * d + jb]) jb = j;
  double vn = 0.0;
  for (int j = 0; j < d; ++j) vn = fma(M[j * d + jb], M[j * d + jb], vn);
  const double vs = vn > 0.0 ? 1.0 / sqrt(vn) : 0.0;
  // w = W0 v, sigma = |w|
  double w = 0.0;
  if (tid < D) {
    for (int j = 0; j < d; ++j) w = fma(bf_w0(s, tr, t, j, tid), M[j * d + jb] * vs, w);
  }
  double part = w * w;
  for (int i = tid + BF_THREADS; i < D; i += BF_THREADS) {
    double wi = 0.0;
    for (int j = 0; j < d; ++j) wi = fma(bf_w0(s, tr, t, j, i), M[j * d + jb] * vs, wi);
    part = fma(wi, wi, part);
  }
  part = wave_sum(part);
  if ((tid & 63) == 0) red[tid >> 6] = part;
  __syncthreads();
  double lam = 0.0;
  for (int q = 0; q < BF_THREADS / 64; ++q) lam += red[q];
  double kl = 0.0, kc = 0.0;
  if (lam > 0.0 && vn > 0.0) {
    const double sigma = sqrt(lam), rs = sqrt(sigma);
    kl = 1.0 / rs;
    kc = rs;
  }
  if (tid < D) bf_put_len(s, tr, t, tid, w * kl);
  for (int i = tid + BF_THREADS; i < D; i += BF_THREADS) {
    double wi = 0.0;
    for (int j = 0; j < d; ++j) wi = fma(bf_w0(s, tr, t, j, i), M[j * d + jb] * vs, wi);
    bf_put_len(s, tr, t, i, wi * kl);
  }
  for (int j = tid; j < d; j += BF_THREADS) bf_put_col(s, tr, t, j, M[j * d + jb] * vs * kc);
}

// The same computation with the d x d products on FP64 MFMA
// (v_mfma_f64_16x16x4f64; d = 16..128).  G and every power of it are
// symmetric, so only the 16 x 16 tiles (ti, tj) with ti <= tj are formed
// (wave w takes the tiles w, w + 4, ... of that list, NT at most) and each
// off-diagonal tile is also stored transposed.  Operand lanes: A[i][k] in
// lane i + 16 k, B[k][j] in lane j + 16 k; output register r of lane l is
// D[l / 16 + 4 r][l % 16].  Both operands of a square are read from rows of
// M (M symmetric: column i of M is row i).  M is row-major with row stride
// d + 1, so the transposed stores of the 16 lanes of a quarter-wave hit 16
// different banks.
typedef double bf_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void bf_store_tile(double* M, int ld, int ti, int tj, int li, int lk, const bf_d4& c,
                                              double f) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const double v = c[r] * f;
    M[(16 * ti + lk + 4 * r) * ld + 16 * tj + li] = v;
    if (ti != tj) M[(16 * tj + li) * ld + 16 * ti + lk + 4 * r] = v;
  }
}

// trace of the d x d matrix M (row stride ld): lane-strided partial sums and
// a xor-butterfly wave sum, which leaves the same bits in every lane of every
// wave, so the stopping test stays uniform over the workgroup
__device__ __forceinline__ double bf_trace(const double* M, int ld, int d, int lane) {
  double v = 0.0;
  for (int j = lane; j < d; j += 64) v += M[j * ld + j];
  return wave_sum(v);
}

template <int NT>
__global__ __launch_bounds__(BF_THREADS) void k_bfly_gram_mfma(BflyStep s) {
  extern __shared__ double lds[];
  __shared__ double red[BF_THREADS / 64];
  const int t = bf_block_index(blockIdx.x, gridDim.x), tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
  const bool tr = s.logBx < s.logBy;
  const int logD = tr ? s.logBy : s.logBx, logd = tr ? s.logBx : s.logBy;
  const int D = 1 << logD, d = 1 << logd, ld = d + 1;
  double* M = lds;  // d x d, row stride ld
  double* B = lds;  // 16 x d staged rows of W0: only live before M is first written
  const int td = d >> 4, nup = td * (td + 1) / 2;
  // this wave's upper tiles
  int tI[NT], tJ[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    int u = wave + 4 * q, ti = 0;
    while (ti < td && u >= td - ti) {
      u -= td - ti;
      ++ti;
    }
    tI[q] = ti;
    tJ[q] = ti + u;
  }
  bf_d4 acc[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) acc[q] = bf_d4{0.0, 0.0, 0.0, 0.0};
  // G = W0^T W0: G[p][q] = sum over the rows r of W0 of B[r][p] B[r][q].
  // Entry e = r d + j of a 16-row chunk is staged by thread e % 256; the next
  // chunk's loads are issued before the current chunk's products.
  constexpr int PF = 16 * 128 / BF_THREADS;  // staged entries per thread at d = 128
  double pre[PF];
#pragma unroll
  for (int k = 0; k < PF; ++k) {
    const int e = tid + k * BF_THREADS;
    if (e < 16 * d) pre[k] = bf_w0(s, tr, t, e & (d - 1), e >> logd);
  }
  for (int i0 = 0; i0 < D; i0 += 16) {
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      const int e = tid + k * BF_THREADS;
      if (e < 16 * d) B[e] = pre[k];
    }
    __syncthreads();
    if (i0 + 16 < D) {
#pragma unroll
      for (int k = 0; k < PF; ++k) {
        const int e = tid + k * BF_THREADS;
        if (e < 16 * d) pre[k] = bf_w0(s, tr, t, e & (d - 1), i0 + 16 + (e >> logd));
      }
    }
    // tiles inner: the wave's NT accumulation chains interleave
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const double* row = B + (4 * ks + lk) * d;
#pragma unroll
      for (int q = 0; q < NT; ++q)
        if (wave + 4 * q < nup)
          acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(row[16 * tI[q] + li], row[16 * tJ[q] + li], acc[q], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < NT; ++q)
    if (wave + 4 * q < nup) bf_store_tile(M, ld, tI[q], tJ[q], li, lk, acc[q], 1.0);
  __syncthreads();
  double tau = bf_trace(M, ld, d, lane);
  if (tau > 0.0) {
    double scale = 1.0 / tau;  // M holds the current matrix / scale
    for (int k = 0; k < BF_MAX_SQ; ++k) {
      const double sc2 = scale * scale;
#pragma unroll
      for (int q = 0; q < NT; ++q) acc[q] = bf_d4{0.0, 0.0, 0.0, 0.0};
      for (int kk = 0; kk < d; kk += 4) {
        const double* row = M + (kk + lk) * ld + li;
#pragma unroll
        for (int q = 0; q < NT; ++q)
          if (wave + 4 * q < nup)
            acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(row[16 * tI[q]], row[16 * tJ[q]], acc[q], 0, 0, 0);
      }
      __syncthreads();  // every tile of the square is formed before M is overwritten
#pragma unroll
      for (int q = 0; q < NT; ++q)
        if (wave + 4 * q < nup) bf_store_tile(M, ld, tI[q], tJ[q], li, lk, acc[q], sc2);
      __syncthreads();
      const double tr2 = bf_trace(M, ld, d, lane);
      scale = 1.0 / tr2;
      if (tr2 >= 1.0 - 16.0 * DBL_EPSILON) break;  // tr(M^2) = 1: M is rank one
    }
  }
  __syncthreads();
  // v: the column of largest diagonal entry, normalised
  int jb = 0;
  for (int j = 1; j < d; ++j)
    if (M[j * ld + j] > M[jb * ld + jb]) jb = j;
  double vn = 0.0;
  for (int j = 0; j < d; ++j) vn = fma(M[j * ld + jb], M[j * ld + jb], vn);
  const double vs = vn > 0.0 ? 1.0 / sqrt(vn) : 0.0;
  // w = W0 v, sigma = |w|; entry i = tid is kept for the store below
  double w = 0.0;
  if (tid < D)
    for (int j = 0; j < d; ++j) w = fma(bf_w0(s, tr, t, j, tid), M[j * ld + jb] * vs, w);
  double part = w * w;
  for (int i = tid + BF_THREADS; i < D; i += BF_THREADS) {
    double wi = 0.0;
    for (int j = 0; j < d; ++j) wi = fma(bf_w0(s, tr, t, j, i), M[j * ld + jb] * vs, wi);
    part = fma(wi, wi, part);
  }
  part = wave_sum(part);
  if (lane == 0) red[wave] = part;
  __syncthreads();
  double lam = 0.0;
  for (int q = 0; q < BF_THREADS / 64; ++q) lam += red[q];
  double kl = 0.0, kc = 0.0;
  if (lam > 0.0 && vn > 0.0) {
    const double sigma = sqrt(lam), rs = sqrt(sigma);
    kl = 1.0 / rs;
    kc = rs;
  }
  if (tid < D) bf_put_len(s, tr, t, tid, w * kl);
  for (int i = tid + BF_THREADS; i < D; i += BF_THREADS) {
    double wi = 0.0;
    for (int j = 0; j < d; ++j) wi = fma(bf_w0(s, tr, t, j, i), M[j * ld + jb] * vs, wi);
    bf_put_len(s, tr, t, i, wi * kl);
  }
  for (int j = tid; j < d; j += BF_THREADS) bf_put_col(s, tr, t, j, M[j * ld + jb] * vs * kc);
}

// The root node (the whole matrix, row-major: value (r, c) at r n + c) from
// the caller's matrix with its columns permuted: root(r, c) = A(r, pinv[c])
// (pinv == nullptr: no permutation).  Column-major A goes through 32 x 32
// LDS tiles so that reads run down columns and writes along rows.
constexpr int BF_TT = 32;
__global__ __launch_bounds__(256) void k_bfly_root_cm(const double* A, size_t lda, int n, const int* pinv,
                                                        double* out) {
  __shared__ double T[BF_TT][BF_TT + 1];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const int r0 = blockIdx.x * BF_TT, c0 = blockIdx.y * BF_TT;
  for (int k = ty; k < BF_TT; k += 8) {
    const int c = c0 + k;
    const int src = pinv ? pinv[c] : c;
    T[k][tx] = A[(size_t)(r0 + tx) + (size_t)src * lda];
  }
  __syncthreads();
  for (int k = ty; k < BF_TT; k += 8) out[(size_t)(r0 + k) * n + c0 + tx] = T[tx][k];
}

__global__ void k_bfly_root_small(const double* A, size_t lda, int n, const int* pinv, double* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int r = e / n, c = e % n;
  out[e] = A[(size_t)r + (size_t)(pinv ? pinv[c] : c) * lda];
}

__global__ void k_bfly_root_rm(const double* A, size_t lda, int n, const int* pinv, double* out) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * n) return;
  const int r = (int)(e / n), c = (int)(e % n);
  out[e] = A[(size_t)r * lda + (pinv ? pinv[c] : c)];
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

// CSR (with its CSC twin and both 8-slot ELL views) of a 2-entry-per-row
// leaf node, built in place: row r holds v[2r] at column c0(r) and v[2r + 1]
// at c0(r) + s, c0(r) = (r & ~(2s - 1)) + (r & (s - 1)).  The support is
// symmetric, so column c (before relabelling) holds rows c0(c) and c0(c) + s,
// and row i's value there is v[2i + ((c & s) != 0)].  With colmap (a
// permutation, colinv its inverse) old column c becomes column colmap[c];
// each row's two entries are then sorted by their new column.
__global__ void k_bfly_leaf_csr(const double* v, int n, int logs, const int* colmap, const int* colinv, int* rp,
                                int* ci, double* cv_row, int* cp, int* ri, double* cv_col, int* ell_ci,
                                double* ell_v, int* tell_ci, double* tell_v) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int s = 1 << logs;
  // row r
  int c0 = (r & ~(2 * s - 1)) + (r & (s - 1)), c1 = c0 + s;
  double v0 = v[2 * (size_t)r], v1 = v[2 * (size_t)r + 1];
  if (colmap) {
    c0 = colmap[c0];
    c1 = colmap[c1];
    if (c1 < c0) {
      const int tc = c0;
      c0 = c1;
      c1 = tc;
      const double tv = v0;
      v0 = v1;
      v1 = tv;
    }
  }
  rp[r] = 2 * r;
  ci[2 * (size_t)r] = c0;
  ci[2 * (size_t)r + 1] = c1;
  cv_row[2 * (size_t)r] = v0;
  cv_row[2 * (size_t)r + 1] = v1;
  // column r (new label)
  const int oc = colinv ? colinv[r] : r;
  const int rb = (oc & ~(2 * s - 1)) + (oc & (s - 1));
  const int slot = (oc & s) ? 1 : 0;
  const double w0 = v[2 * (size_t)rb + slot], w1 = v[2 * (size_t)(rb + s) + slot];
  cp[r] = 2 * r;
  ri[2 * (size_t)r] = rb;
  ri[2 * (size_t)r + 1] = rb + s;
  cv_col[2 * (size_t)r] = w0;
  cv_col[2 * (size_t)r + 1] = w1;
  for (int q = 0; q < 8; ++q) {
    const size_t e = (size_t)r * 8 + q;
    ell_ci[e] = q == 0 ? c0 : q == 1 ? c1 : -1;
    ell_v[e] = q == 0 ? v0 : q == 1 ? v1 : 0.0;
    tell_ci[e] = q == 0 ? rb : q == 1 ? rb + s : -1;
    tell_v[e] = q == 0 ? w0 : q == 1 ? w1 : 0.0;
  }
  if (r == 0) {
    rp[n] = 2 * n;
    cp[n] = 2 * n;
  }
}

void bfly_leaf_csr(const double* v, int n, int logs, const int* colmap, const int* colinv, DMat& M, hipStream_t st) {
  FH_CHECK(n >= 4 && (n & (n - 1)) == 0 && logs >= 0 && (2 << logs) <= n, "butterfly: bad leaf node");
  FH_CHECK((colmap == nullptr) == (colinv == nullptr), "butterfly: leaf relabelling needs both maps");
  const size_t nz = 2 * (size_t)n, ne = 8 * (size_t)n;
  M.kind = K_CSR;
  M.dtype = FH_F64;
  M.m = M.n = n;
  M.nnz = (int64_t)nz;
  M.rp.alloc(sizeof(int) * (n + 1));
  M.ci.alloc(sizeof(int) * nz);
  M.v.alloc(sizeof(double) * nz);
  M.cp.alloc(sizeof(int) * (n + 1));
  M.ri.alloc(sizeof(int) * nz);
  M.cv.alloc(sizeof(double) * nz);
  M.has_csc = true;
  M.max_row = M.max_col = 2;
  M.csr_ell_w = M.csr_tell_w = 8;  // the FP64 gather kernels' slot quantum
  M.csr_ell_ci.alloc(sizeof(int) * ne);
  M.csr_ell_v.alloc(sizeof(double) * ne);
  M.csr_tell_ci.alloc(sizeof(int) * ne);
  M.csr_tell_v.alloc(sizeof(double) * ne);
  hipLaunchKernelGGL(k_bfly_leaf_csr, dim3((n + 255) / 256), dim3(256), 0, st, v, n, logs, colmap, colinv,
                     M.rp.as<int>(), M.ci.as<int>(), M.v.as<double>(), M.cp.as<int>(), M.ri.as<int>(),
                     M.cv.as<double>(), M.csr_ell_ci.as<int>(), M.csr_ell_v.as<double>(), M.csr_tell_ci.as<int>(),
                     M.csr_tell_v.as<double>());
  FH_LAUNCH_CHECK();
}

static size_t bfly_lds_budget() {
  static size_t budget = 0;
  if (budget) return budget;
  int dev = 0, v = 0;
  size_t b = 64 * 1024 - 256;  // the kernel's static LDS (wave sums) comes on top
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 64 * 1024) {
    const int want = v - 1024;
    const void* fns[] = {(const void*)k_bfly_gram<4>, (const void*)k_bfly_gram<8>,
                         (const void*)k_bfly_gram_mfma<1>, (const void*)k_bfly_gram_mfma<3>,
                         (const void*)k_bfly_gram_mfma<9>};
    bool ok = true;
    for (const void* f : fns)
      ok = ok && hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, want) == hipSuccess;
    if (ok) b = (size_t)want;
  }
  (void)hipGetLastError();  // a refused attribute leaves the 64 KiB budget, not an error
  budget = b;
  return budget;
}

// d x d products on MFMA for d >= 16 (FH_BFLY_MFMA=0: the VALU kernel)
static bool bfly_use_mfma(int logd) {
  const char* e = getenv("FH_BFLY_MFMA");  // read per split: a caller may switch between factorizations
  return !(e && e[0] == '0') && logd >= 4;
}

size_t bfly_block_lds_bytes(int logBx, int logBy) {
  const int logd = std::min(logBx, logBy);
  const size_t d = (size_t)1 << logd, D = (size_t)1 << std::max(logBx, logBy);
  if (bfly_use_mfma(logd)) return sizeof(double) * d * (d + 1);  // padded M (the staged rows share it)
  return sizeof(double) * (d * d + std::min<size_t>(BF_CHUNK, D) * d);  // M and the staged rows
}

bool bfly_block_fits(int logBx, int logBy) {
  const int logd = std::min(logBx, logBy);
  if (logd <= 1) return true;   // closed form, any length
  if (logd > 7) return false;   // one 8 x 8 tile of d x d per thread: d <= 128
  return bfly_block_lds_bytes(logBx, logBy) <= bfly_lds_budget();
}

void bfly_split(const BflyStep& s, hipStream_t st) {
  FH_CHECK(s.n >= 4 && (s.n & (s.n - 1)) == 0, "butterfly: order must be a power of two >= 4");
  FH_CHECK(s.logBx >= 1 && s.logBy >= 1, "butterfly: empty split");
  FH_CHECK(s.logBx + s.logsx == s.in_logB + s.in_logs && s.logsy == s.in_logs && s.logBy + s.logsy == s.logsx,
           "butterfly: split does not match its node");
  if (std::min(s.logBx, s.logBy) == 1) {
    const int logD = std::max(s.logBx, s.logBy);
    const char* e = getenv("FH_BFLY_D2_WIDE");  // 0: one wave per block at every length
    if (!(e && e[0] == '0') && logD >= 11 && logD <= 13) {
      if (logD == 11)
        hipLaunchKernelGGL(k_bfly_d2_wide<8>, dim3(s.n), dim3(BF_THREADS), 0, st, s);
      else if (logD == 12)
        hipLaunchKernelGGL(k_bfly_d2_wide<16>, dim3(s.n), dim3(BF_THREADS), 0, st, s);
      else
        hipLaunchKernelGGL(k_bfly_d2_wide<32>, dim3(s.n), dim3(BF_THREADS), 0, st, s);
      FH_LAUNCH_CHECK();
      return;
    }
    const int per = BF_THREADS / 64;
    hipLaunchKernelGGL(k_bfly_d2, dim3((s.n + per - 1) / per), dim3(BF_THREADS), 0, st, s);
  } else {
    FH_CHECK(bfly_block_fits(s.logBx, s.logBy), "butterfly: block too large for one workgroup's LDS");
    const size_t lds = bfly_block_lds_bytes(s.logBx, s.logBy);
    const int logd = std::min(s.logBx, s.logBy);
    if (bfly_use_mfma(logd)) {
      // upper 16 x 16 tiles per wave: 1 (d <= 32), 3 (d = 64), 9 (d = 128)
      if (logd <= 5)
        hipLaunchKernelGGL(k_bfly_gram_mfma<1>, dim3(s.n), dim3(BF_THREADS), lds, st, s);
      else if (logd == 6)
        hipLaunchKernelGGL(k_bfly_gram_mfma<3>, dim3(s.n), dim3(BF_THREADS), lds, st, s);
      else
        hipLaunchKernelGGL(k_bfly_gram_mfma<9>, dim3(s.n), dim3(BF_THREADS), lds, st, s);
    } else if (logd <= 6)
      hipLaunchKernelGGL(k_bfly_gram<4>, dim3(s.n), dim3(BF_THREADS), lds, st, s);
    else
      hipLaunchKernelGGL(k_bfly_gram<8>, dim3(s.n), dim3(BF_THREADS), lds, st, s);
  }
  FH_LAUNCH_CHECK();
}

void bfly_root(const double* A, size_t lda, int row_major, int n, const int* pinv_dev, double* out, hipStream_t st) {
  FH_CHECK(n >= 4 && (n & (n - 1)) == 0, "butterfly: order must be a power of two >= 4");
  if (row_major || n < BF_TT) {
    FH_CHECK(row_major || lda >= (size_t)n, "butterfly: bad leading dimension");
    if (!row_major) {
      // small column-major matrices: one thread per entry, reading A(r, c) at r + c lda
      hipLaunchKernelGGL(k_bfly_root_small, dim3((n * n + 255) / 256), dim3(256), 0, st, A, lda, n, pinv_dev, out);
    } else {
      const size_t nn = (size_t)n * n;
      hipLaunchKernelGGL(k_bfly_root_rm, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, A, lda, n, pinv_dev,
                         out);
    }
  } else {
    hipLaunchKernelGGL(k_bfly_root_cm, dim3(n / BF_TT, n / BF_TT), dim3(256), 0, st, A, lda, n, pinv_dev, out);
  }
  FH_LAUNCH_CHECK();
}

void bfly_scatter_dense(const double* v, int n, int logs, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_bfly_scatter, dim3((n + 255) / 256), dim3(256), 0, st, v, n, logs, out);
  FH_LAUNCH_CHECK();
}

}  // namespace fh
