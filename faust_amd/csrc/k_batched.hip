// k_batched.hip -- batched PALM4MSA: many independent factorizations of
// uneven sizes advanced in lock-step, FP32 or FP64.
//
// Every per-target scalar (lambda, the two spectral norms of the current
// factor update) lives in device memory and every kernel reads it there, so
// the kernel sequence of one PALM iteration is fixed and is captured once as
// a HIP graph, then replayed for all iterations: no host round trip inside
// the loop.  Each batched kernel walks a flat work list (target, tile), so
// targets of 64..512 rows share one launch.
//
// Algorithm per target: palm4msa2 with packing_RL and left-to-right updates
// (faust_palm4msa2020.hpp:47-300 with the ParamsPalm4MSA defaults), all
// factors square and kept dense.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "fh_api_common.hpp"
#include "fh_batched.hpp"

namespace fh {

// ---------------------------------------------------------------- grouped GEMM
// C_out = alpha_b * op(A) op(B) + beta * C_in, alpha_b from the per-target
// scalars: mode 0 -> 1, 1 -> lambda_b, 2 -> -lambda_b / c_b with
// c_b = 1.001 lambda_b^2 nR_b^2 nL_b^2 (update_fact's Lipschitz step).

template <class T>
__device__ __forceinline__ T job_alpha(int mode, int b, const BScal& s) {
  if (mode == 0) return T(1);
  const double lam = s.lambda[b];
  if (mode == 1) return (T)lam;
  const double nr = s.nR[b], nl = s.nL[b];
  const double c = 1.001 * lam * lam * nr * nr * nl * nl;
  // update_fact throws when a 2-norm is zero or NaN; a batched target whose
  // step is not finite and positive keeps its factor instead
  if (!(c > 0.0) || !isfinite(c) || !isfinite(lam)) return T(0);
  const T a = (T)(-lam / c);
  return isfinite(a) ? a : T(0);  // beyond the FP32 range: no step rather than inf
}

// Workgroup -> tile.  The hardware deals workgroups to the 8 XCDs in turn
// (workgroup b to XCD b % 8); with per = ceil(ntiles / 8), XCD x gets the
// contiguous tiles [x per, (x + 1) per), so that the tiles of one job, which
// read the same operands, mostly share one XCD's L2.  per = 0 keeps tile =
// workgroup.  Workgroups past the last tile return at once.
__device__ __forceinline__ int xcd_tile(int per) {
  const int b = blockIdx.x;
  return per > 0 ? (b & 7) * per + (b >> 3) : b;
}

template <class T, int TILE>
__global__ __launch_bounds__(256) void k_bgemm(const BGemmJob<T>* __restrict__ jobs, const int3* __restrict__ tiles,
                                               BScal sc, int ntiles, int per) {
  constexpr int MT = TILE / 16;  // micro-tile edge per thread
  constexpr int KS = 16;
  __shared__ T As[KS][TILE + 1];
  __shared__ T Bs[KS][TILE + 1];
  const int tix = xcd_tile(per);
  if (tix >= ntiles) return;
  const int3 tl = tiles[tix];
  const BGemmJob<T> J = jobs[tl.x];
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const int i0 = tl.y * TILE, j0 = tl.z * TILE;
  T acc[MT][MT];
#pragma unroll
  for (int r = 0; r < MT; ++r)
#pragma unroll
    for (int c = 0; c < MT; ++c) acc[r][c] = T(0);
  for (int k0 = 0; k0 < J.k; k0 += KS) {
    for (int e = t; e < TILE * KS; e += 256) {
      int ii, kk;
      if (!J.transA) { ii = e % TILE; kk = e / TILE; } else { kk = e % KS; ii = e / KS; }
      const int gi = i0 + ii, gk = k0 + kk;
      As[kk][ii] = (gi < J.m && gk < J.k) ? (J.transA ? J.A[(size_t)gi * J.k + gk] : J.A[(size_t)gk * J.m + gi])
                                          : T(0);
    }
    for (int e = t; e < TILE * KS; e += 256) {
      int jj, kk;
      if (J.transB) { jj = e % TILE; kk = e / TILE; } else { kk = e % KS; jj = e / KS; }
      const int gj = j0 + jj, gk = k0 + kk;
      Bs[kk][jj] = (gj < J.n && gk < J.k) ? (J.transB ? J.B[(size_t)gk * J.n + gj] : J.B[(size_t)gj * J.k + gk])
                                          : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) {
      T a[MT], b[MT];
#pragma unroll
      for (int r = 0; r < MT; ++r) a[r] = As[kk][ty + 16 * r];
#pragma unroll
      for (int c = 0; c < MT; ++c) b[c] = Bs[kk][tx + 16 * c];
#pragma unroll
      for (int r = 0; r < MT; ++r)
#pragma unroll
        for (int c = 0; c < MT; ++c) acc[r][c] = fma(a[r], b[c], acc[r][c]);
    }
    __syncthreads();
  }
  const T alpha = job_alpha<T>(J.alpha_mode, J.target, sc);
  const T beta = J.beta_mode == 0 ? T(0) : (J.beta_mode == 1 ? T(-1) : T(1));
#pragma unroll
  for (int r = 0; r < MT; ++r) {
    const int gi = i0 + ty + 16 * r;
    if (gi >= J.m) continue;
#pragma unroll
    for (int c = 0; c < MT; ++c) {
      const int gj = j0 + tx + 16 * c;
      if (gj >= J.n) continue;
      const size_t o = (size_t)gj * J.m + gi;
      T v = alpha * acc[r][c];
      if (J.beta_mode) v = fma(beta, J.Cin[o], v);
      J.Cout[o] = v;
    }
  }
}

// FP32 on the matrix cores: v_mfma_f32_32x32x2_f32, workgroup tile 64 x 64
// (four waves of 32 x 32), K staged through LDS 32 at a time.  Operand
// lanes: A lane l holds op(A)[row l % 32][k + l / 32], B lane l holds
// op(B)[k + l / 32][col l % 32]; the accumulator holds rows
// 8*(r/4) + 4*(l/32) + r%4 of column l % 32 (r = 0..15).
typedef float f32x16 __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(256) void k_bgemm_mfma_f32(const BGemmJob<float>* __restrict__ jobs,
                                                        const int3* __restrict__ tiles, BScal sc, int ntiles,
                                                        int per) {
  constexpr int TILE = 64, KS = 32;
  __shared__ float As[KS][TILE + 4];
  __shared__ float Bs[KS][TILE + 4];
  const int tix = xcd_tile(per);
  if (tix >= ntiles) return;
  const int3 tl = tiles[tix];
  const BGemmJob<float> J = jobs[tl.x];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wr = wave & 1, wc = wave >> 1;
  const int i0 = tl.y * TILE, j0 = tl.z * TILE;
  f32x16 acc = {};
  const bool inside = i0 + TILE <= J.m && j0 + TILE <= J.n;
  for (int k0 = 0; k0 < J.k; k0 += KS) {
    if (inside && k0 + KS <= J.k) {
      // whole tile inside: 8 + 8 unconditional loads per thread, all in flight
      float av[TILE * KS / 256], bv[TILE * KS / 256];
#pragma unroll
      for (int q = 0; q < TILE * KS / 256; ++q) {
        const int e = t + 256 * q;
        const int ii = J.transA ? e / KS : e % TILE, ka = J.transA ? e % KS : e / TILE;
        av[q] = J.transA ? J.A[(size_t)(i0 + ii) * J.k + k0 + ka] : J.A[(size_t)(k0 + ka) * J.m + i0 + ii];
        const int jj = J.transB ? e % TILE : e / KS, kb = J.transB ? e / TILE : e % KS;
        bv[q] = J.transB ? J.B[(size_t)(k0 + kb) * J.n + j0 + jj] : J.B[(size_t)(j0 + jj) * J.k + k0 + kb];
      }
#pragma unroll
      for (int q = 0; q < TILE * KS / 256; ++q) {
        const int e = t + 256 * q;
        const int ii = J.transA ? e / KS : e % TILE, ka = J.transA ? e % KS : e / TILE;
        As[ka][ii] = av[q];
        const int jj = J.transB ? e % TILE : e / KS, kb = J.transB ? e / TILE : e % KS;
        Bs[kb][jj] = bv[q];
      }
    } else {
      for (int e = t; e < TILE * KS; e += 256) {
        int ii, kk;
        if (!J.transA) { ii = e % TILE; kk = e / TILE; } else { kk = e % KS; ii = e / KS; }
        const int gi = i0 + ii, gk = k0 + kk;
        As[kk][ii] = (gi < J.m && gk < J.k) ? (J.transA ? J.A[(size_t)gi * J.k + gk] : J.A[(size_t)gk * J.m + gi])
                                            : 0.0f;
      }
      for (int e = t; e < TILE * KS; e += 256) {
        int jj, kk;
        if (J.transB) { jj = e % TILE; kk = e / TILE; } else { kk = e % KS; jj = e / KS; }
        const int gj = j0 + jj, gk = k0 + kk;
        Bs[kk][jj] = (gj < J.n && gk < J.k) ? (J.transB ? J.B[(size_t)gk * J.n + gj] : J.B[(size_t)gj * J.k + gk])
                                            : 0.0f;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KS; kk += 2) {
      const float a = As[kk + (lane >> 5)][wr * 32 + (lane & 31)];
      const float b = Bs[kk + (lane >> 5)][wc * 32 + (lane & 31)];
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  const float alpha = job_alpha<float>(J.alpha_mode, J.target, sc);
  const float beta = J.beta_mode == 0 ? 0.0f : (J.beta_mode == 1 ? -1.0f : 1.0f);
  const int gj = j0 + wc * 32 + (lane & 31);
  if (gj >= J.n) return;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int gi = i0 + wr * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (gi >= J.m) continue;
    const size_t o = (size_t)gj * J.m + gi;
    float v = alpha * acc[r];
    if (J.beta_mode) v = fmaf(beta, J.Cin[o], v);
    J.Cout[o] = v;
  }
}

// ---------------------------------------------------------------- power iteration
// One workgroup per job: lambda_max(P P^T) by power iteration with the stop
// rule of power_iteration (faust_linear_algebra.hpp:846-889); the result
// sqrt(lambda) goes to out[target].  warm: start from (and store) the last
// normalised vector of this job slot.

__device__ double block_sum_d(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s >= 1; s >>= 1) {
    if (t < s) red[t] = red[t] + red[t + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// The step size stays FP64 whatever the factor precision: vectors and sums
// are double, matrix entries are widened on load.
template <class T>
__global__ __launch_bounds__(256) void k_bpower(const BPowJob<T>* __restrict__ jobs, int maxit, double thr,
                                                int warm) {
  __shared__ double x[BATCH_NMAX], y[BATCH_NMAX];
  __shared__ double red[256];
  const BPowJob<T> J = jobs[blockIdx.x];
  const int n = J.n, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (J.P == nullptr) {
    if (t == 0) J.out[J.target] = 1.0;
    return;
  }
  for (int i = t; i < n; i += 256) x[i] = J.vec[i];
  __syncthreads();
  double lambda_old = 1.0, lambda = 0.0;
  int it = 0;
  bool zero = false;
  while ((fabs(lambda_old - lambda) > fabs(thr) || fabs(lambda) <= fabs(thr)) && it < maxit) {
    ++it;
    lambda_old = lambda;
    double ss = 0.0;
    for (int i = t; i < n; i += 256) ss = fma(x[i], x[i], ss);
    const double nrm = sqrt(block_sum_d(ss, red));
    if (nrm == 0.0) {
      zero = true;
      break;
    }
    const double inv = 1.0 / nrm;
    for (int i = t; i < n; i += 256) x[i] = x[i] * inv;
    __syncthreads();
    // y = P^T x: one wave per column of P (contiguous), lanes split k
    for (int i = wave; i < n; i += 4) {
      const T* col = J.P + (size_t)i * n;
      double a = 0.0;
      for (int k = lane; k < n; k += 64) a = fma((double)col[k], x[k], a);
      a = wave_sum_d(a);
      if (lane == 0) y[i] = a;
    }
    __syncthreads();
    // z = P y (threads over rows: coalesced), lambda = <x, z>
    double dot = 0.0;
    double z[BATCH_NMAX / 256];
    int nz = 0;
#pragma unroll
    for (int q = 0; q < BATCH_NMAX / 256; ++q) {
      const int i = t + 256 * q;
      double a = 0.0;
      if (i < n) {
        for (int k = 0; k < n; ++k) a = fma((double)J.P[(size_t)k * n + i], y[k], a);
        dot = fma(x[i], a, dot);
        nz |= (a != 0.0);
      }
      z[q] = a;
    }
    lambda = block_sum_d(dot, red);
    if (!__syncthreads_or(nz)) {
      lambda = 0.0;
      zero = true;
      break;
    }
#pragma unroll
    for (int q = 0; q < BATCH_NMAX / 256; ++q)
      if (t + 256 * q < n) x[t + 256 * q] = z[q];
    __syncthreads();
  }
  if (warm && !zero)
    for (int i = t; i < n; i += 256) J.vec[i] = x[i];
  if (t == 0) J.out[J.target] = sqrt(fabs(lambda));
}

// ---------------------------------------------------------------- sparse factors
// Compressed rows and columns of a dense n x n factor (one 256-thread
// workgroup per job), entries of a row in column order and of a column in row
// order whatever the thread timing.  Entries beyond cap are not written (the
// caller sizes cap to the projection's nonzero count).
__device__ void block_excl_scan_int(int* a, int n, int* part) {
  // a[0..n) -> exclusive prefix sums; part: 256 ints of LDS scratch
  const int t = threadIdx.x;
  const int per = (n + 255) / 256;
  const int lo = min(t * per, n), hi = min(lo + per, n);
  int s = 0;
  for (int i = lo; i < hi; ++i) s += a[i];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int i = 0; i < 256; ++i) {
      const int v = part[i];
      part[i] = run;
      run += v;
    }
  }
  __syncthreads();
  int run = part[t];
  for (int i = lo; i < hi; ++i) {
    const int v = a[i];
    a[i] = run;
    run += v;
  }
  __syncthreads();
}

// From the projection's list of kept entries (at most BSP_SORT): the list
// sorted in LDS by column-major position gives the compressed columns in
// order (entry q of the sorted list is entry q of the columns), sorted again
// by row-major position the compressed rows; the pointers are lower bounds
// in the sorted lists.  The result equals the dense scan's.
constexpr int BSP_SORT = 8192;

__device__ void bitonic_sort_lds(int* a, int P) {  // P a power of two, 256 threads
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += 256) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & k) == 0;
          const int x = a[i], y = a[l];
          if ((x > y) == up) {
            a[i] = y;
            a[l] = x;
          }
        }
      }
      __syncthreads();
    }
}

__device__ __forceinline__ int lower_bound_lds(const int* a, int cnt, int key) {
  int lo = 0, hi = cnt;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

template <class T>
__device__ void bsparse_from_list(const BSpBuildJob<T>& J, int nk, int* keys) {
  const int n = J.n, t = threadIdx.x;
  const BSpFac<T> F = J.F;
  int P = 1;
  while (P < nk) P <<= 1;
  for (int q = t; q < P; q += 256) keys[q] = q < nk ? F.klist[q] : INT_MAX;
  __syncthreads();
  bitonic_sort_lds(keys, P);
  // compressed columns: positions j * n + i ascending
  for (int j = t; j <= n; j += 256) F.cp[j] = lower_bound_lds(keys, nk, j * n);
  for (int q = t; q < nk; q += 256) {
    const int idx = keys[q], j = idx / n;
    F.ri[q] = idx - j * n;
    F.cv[q] = J.M[idx];
  }
  __syncthreads();
  // compressed rows: positions i * n + j ascending
  for (int q = t; q < nk; q += 256) {
    const int idx = keys[q], j = idx / n, i = idx - j * n;
    keys[q] = i * n + j;
  }
  __syncthreads();
  bitonic_sort_lds(keys, P);
  for (int i = t; i <= n; i += 256) F.rp[i] = lower_bound_lds(keys, nk, i * n);
  for (int q = t; q < nk; q += 256) {
    const int key = keys[q], i = key / n, j = key - i * n;
    F.ci[q] = j;
    F.v[q] = J.M[(size_t)j * n + i];
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_bsparse(const BSpBuildJob<T>* __restrict__ jobs) {
  __shared__ int cc[BATCH_NMAX + 1], rc[BATCH_NMAX + 1];
  __shared__ int part[256];
  __shared__ int keys[BSP_SORT];
  const BSpBuildJob<T> J = jobs[blockIdx.x];
  const int n = J.n, t = threadIdx.x;
  const BSpFac<T> F = J.F;
  {
    const int nk = F.kcount ? *F.kcount : -1;
    if (nk >= 0 && nk <= F.cap && nk <= BSP_SORT) {
      bsparse_from_list(J, nk, keys);
      return;
    }
  }
  // global (not flat) loads, so that their waits do not also wait for the LDS atomics
  const __attribute__((address_space(1))) T* M = (const __attribute__((address_space(1))) T*)J.M;
  for (int i = t; i <= n; i += 256) cc[i] = 0;
  __syncthreads();
  // counts: thread t walks rows t, t + 256, ... column by column (the 256
  // rows of one column step are one coalesced read); row counts in a
  // register, column counts by LDS atomics (integer sums)
  constexpr int UB = 16;  // columns loaded together before they are tested
  for (int i0 = 0; i0 < n; i0 += 256) {
    const int i = i0 + t;
    int c = 0;
    if (i < n)
      for (int j0 = 0; j0 < n; j0 += UB) {
        T xv[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) xv[u] = M[(size_t)min(j0 + u, n - 1) * n + i];
#pragma unroll
        for (int u = 0; u < UB; ++u)
          if (j0 + u < n && xv[u] != T(0)) {
            ++c;
            atomicAdd(&cc[j0 + u], 1);
          }
      }
    if (i < n) rc[i] = c;
  }
  if (t == 0) rc[n] = 0;
  __syncthreads();
  block_excl_scan_int(cc, n + 1, part);
  block_excl_scan_int(rc, n + 1, part);
  // pointers clamped to cap, so that a factor with more entries than room
  // (which a sp projection does not produce) is cut, never read past its arrays
  for (int i = t; i <= n; i += 256) {
    F.cp[i] = min(cc[i], F.cap);
    F.rp[i] = min(rc[i], F.cap);
  }
  __syncthreads();
  // fill: a row's entries in column order from the thread's own walk; a
  // column's entries at slots claimed by LDS atomics (cc becomes the running
  // end of each column), put in row order below
  for (int i0 = 0; i0 < n; i0 += 256) {
    const int i = i0 + t;
    if (i >= n) continue;
    int q = rc[i];
    for (int j0 = 0; j0 < n; j0 += UB) {
      T xv[UB];
#pragma unroll
      for (int u = 0; u < UB; ++u) xv[u] = M[(size_t)min(j0 + u, n - 1) * n + i];
#pragma unroll
      for (int u = 0; u < UB; ++u) {
        const int j = j0 + u;
        const T x = xv[u];
        if (j < n && x != T(0)) {
          if (q < F.cap) {
            F.ci[q] = j;
            F.v[q] = x;
          }
          ++q;
          const int s = atomicAdd(&cc[j], 1);
          if (s < F.cap) {
            F.ri[s] = i;
            F.cv[s] = x;
          }
        }
      }
    }
  }
  __syncthreads();
  // each column's entries by row (insertion sort: a sp factor has a few per column)
  for (int j = t; j < n; j += 256) {
    const int a = min(F.cp[j], F.cap), e = min(cc[j], F.cap);
    for (int p = a + 1; p < e; ++p) {
      const int r = F.ri[p];
      const T x = F.cv[p];
      int q = p - 1;
      while (q >= a && F.ri[q] > r) {
        F.ri[q + 1] = F.ri[q];
        F.cv[q + 1] = F.cv[q];
        --q;
      }
      F.ri[q + 1] = r;
      F.cv[q + 1] = x;
    }
  }
}

// Power iteration on P = F_0 ... F_{m-1} through its sparse factors (one
// 256-thread workgroup per job), the rounds of k_bpower: normalise x, y =
// P^T x (F_0^T first, by compressed columns), z = P y (F_{m-1} first, by
// compressed rows), lambda = <x, z>.  A product of m sparse factors costs
// their entries per round instead of the n^2 of the dense product.  Vectors
// and sums are double, entries widened on load.  Each output entry is one
// fma chain over its row's entries in stored order, in every variant below.
template <class T>
__device__ __forceinline__ void bsp_apply(const BSpFac<T>& F, bool trans, const double* in, double* out, int n) {
  typedef const __attribute__((address_space(1))) int* GI;
  typedef const __attribute__((address_space(1))) T* GV;
  const GI p0 = (GI)(trans ? F.cp : F.rp);
  const GI ix = (GI)(trans ? F.ri : F.ci);
  const GV vv = (GV)(trans ? F.cv : F.v);
  for (int i = threadIdx.x; i < n; i += 256) {
    const int a = p0[i], b = p0[i + 1];
    double acc = 0.0;
    // entries four at a time: their loads issued together (indices clamped
    // inside the factor's arrays), then the fma chain in order
    for (int q0 = a; q0 < b; q0 += 4) {
      int k[4];
      T v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int q = min(q0 + u, b - 1);
        k[u] = ix[q];
        v[u] = vv[q];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (q0 + u < b) acc = fma((double)v[u], in[k[u]], acc);
    }
    out[i] = acc;
  }
  __syncthreads();
}

// A factor descriptor through the scalar cache (it is uniform over the
// workgroup); only the fields the products read.
template <class T>
__device__ __forceinline__ BSpFac<T> bsp_load(const __attribute__((address_space(4))) BSpFac<T>* p) {
  BSpFac<T> f;
  f.rp = p->rp;
  f.ci = p->ci;
  f.v = p->v;
  f.cp = p->cp;
  f.ri = p->ri;
  f.cv = p->cv;
  f.cap = p->cap;
  f.klist = nullptr;
  f.kcount = nullptr;
  return f;
}

// Rows t and t + 256 (n <= 512) of one factor in one orientation: their
// first BSP_REG entries, loaded once per job and kept in registers over all
// the rounds; the rest of a longer row, entries [s, e) of the factor's
// arrays, is read from memory.  A row with fewer entries is padded with
// (-0) x in[BSP_PAD], where the slot BSP_PAD of every vector holds +0:
// -0 x +0 = -0 and acc + (-0) == acc for every acc, so the padding leaves
// each sum bit-identical.
constexpr int BSP_REG = 4;
constexpr int BSP_PAD = 512;
template <class T>
struct BspRows {
  int s[2], e[2];
  int k[2][BSP_REG];
  T v[2][BSP_REG];
  const __attribute__((address_space(1))) int* ix;
  const __attribute__((address_space(1))) T* vv;
};

template <class T>
__device__ __forceinline__ void bsp_rows_load(BspRows<T>& R, const BSpFac<T>& F, bool trans, int n) {
  typedef const __attribute__((address_space(1))) int* GI;
  typedef const __attribute__((address_space(1))) T* GV;
  const GI p0 = (GI)(trans ? F.cp : F.rp);
  R.ix = (GI)(trans ? F.ri : F.ci);
  R.vv = (GV)(trans ? F.cv : F.v);
  int a[2], b[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int i = threadIdx.x + 256 * r;
    a[r] = i < n ? p0[i] : 0;
    b[r] = i < n ? p0[i + 1] : 0;
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int u = 0; u < BSP_REG; ++u) {
      const int q = max(min(a[r] + u, b[r] - 1), 0);
      R.k[r][u] = R.ix[q];
      R.v[r][u] = R.vv[q];
    }
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int u = 0; u < BSP_REG; ++u)
      if (a[r] + u >= b[r]) {
        R.k[r][u] = BSP_PAD;
        R.v[r][u] = -T(0);
      }
    R.s[r] = a[r] + BSP_REG;
    R.e[r] = max(b[r], a[r] + BSP_REG);
  }
}

template <class T>
__device__ __forceinline__ void bsp_rows_apply(const BspRows<T>& R, const double* in, double* out, int n) {
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int i = threadIdx.x + 256 * r;
    if (i < n) {
      double acc = 0.0;
#pragma unroll
      for (int u = 0; u < BSP_REG; ++u) {
        // an empty asm makes the value opaque here, so that its widening is
        // done at each use instead of being hoisted out of the rounds (the
        // widened copies would double the registers the rows take)
        T vu = R.v[r][u];
        asm volatile("" : "+v"(vu));
        acc = fma((double)vu, in[R.k[r][u]], acc);
      }
      if (R.e[r] > R.s[r]) {
        // the rest of a long row; its start is made opaque so that its
        // addresses are formed here rather than kept over the rounds
        int q = R.s[r];
        asm volatile("" : "+v"(q));
        for (; q < R.e[r]; ++q) acc = fma((double)R.vv[q], in[R.ix[q]], acc);
      }
      out[i] = acc;
    }
  }
  __syncthreads();
}

// block sum for the sparse power iteration: lane shuffles inside each wave,
// then the four wave sums in order (a fixed order: the same result every run)
__device__ __forceinline__ double block_sum_w(double v, double* red) {
  v = wave_sum_d(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return r;
}

// two block sums at once (each in the order of block_sum_w); red: 8 doubles
__device__ __forceinline__ void block_sum_w2(double& a, double& b, double* red) {
  a = wave_sum_d(a);
  b = wave_sum_d(b);
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = a;
    red[4 + (threadIdx.x >> 6)] = b;
  }
  __syncthreads();
  a = (red[0] + red[1]) + (red[2] + red[3]);
  b = (red[4] + red[5]) + (red[6] + red[7]);
  __syncthreads();
}

// The rounds of one job.  MC > 0: m == MC factors of edge n <= 512 with their
// rows in registers (cols[c] for P^T, rows[c] for P); MC == 0: any m, rows
// read from memory each round.
template <class T, int MC>
__device__ __forceinline__ void bpower_sp_job(const BPowSpJob<T>& J, int maxit, double thr, int warm, double* x,
                                              double* u, double* w, double* red) {
  const int n = J.n, t = threadIdx.x, m = MC > 0 ? MC : J.m;
  const __attribute__((address_space(4))) BSpFac<T>* fac = (const __attribute__((address_space(4))) BSpFac<T>*)J.fac;
  BspRows<T> cols[MC > 0 ? MC : 1], rows[MC > 0 ? MC : 1];
  if (MC > 0) {
#pragma unroll
    for (int c = 0; c < (MC > 0 ? MC : 1); ++c) {
      const BSpFac<T> F = bsp_load(fac + c);
      bsp_rows_load(cols[c], F, true, n);
      bsp_rows_load(rows[c], F, false, n);
    }
  }
  for (int i = t; i < n; i += 256) x[i] = J.vec[i];
  __syncthreads();
  // Each round: x = src / ||src||, z = P P^T x, lambda = <x, z>, and the
  // next round's src = z.  The sum of squares of z is taken in the pass of
  // the dot product and reduced with it, and z is normalised straight from
  // its buffer: the same sums in the same order as normalising a copy of z
  // at the start of the next round.
  double sstot;
  {
    double ss = 0.0;
    for (int i = t; i < n; i += 256) ss = fma(x[i], x[i], ss);
    sstot = block_sum_w(ss, red);
  }
  const double* src = x;
  double lambda_old = 1.0, lambda = 0.0;
  int it = 0;
  bool zero = false;
  while ((fabs(lambda_old - lambda) > fabs(thr) || fabs(lambda) <= fabs(thr)) && it < maxit) {
    ++it;
    lambda_old = lambda;
    const double nrm = sqrt(sstot);
    if (nrm == 0.0) {
      zero = true;
      break;
    }
    const double inv = 1.0 / nrm;
    for (int i = t; i < n; i += 256) x[i] = src[i] * inv;
    __syncthreads();
    const double* cur = x;
    double* nxt = u;
    if (MC > 0) {
      // y = P^T x = F_{m-1}^T ... F_0^T x, then z = P y = F_0 ... F_{m-1} y
#pragma unroll
      for (int c = 0; c < (MC > 0 ? MC : 1); ++c) {
        bsp_rows_apply(cols[c], cur, nxt, n);
        cur = nxt;
        nxt = (nxt == u) ? w : u;
      }
#pragma unroll
      for (int c = (MC > 0 ? MC : 1) - 1; c >= 0; --c) {
        bsp_rows_apply(rows[c], cur, nxt, n);
        cur = nxt;
        nxt = (nxt == u) ? w : u;
      }
    } else {
      for (int c = 0; c < m; ++c) {
        bsp_apply<T>(bsp_load(fac + c), true, cur, nxt, n);
        cur = nxt;
        nxt = (nxt == u) ? w : u;
      }
      for (int c = m - 1; c >= 0; --c) {
        bsp_apply<T>(bsp_load(fac + c), false, cur, nxt, n);
        cur = nxt;
        nxt = (nxt == u) ? w : u;
      }
    }
    double dot = 0.0, ss = 0.0;
    int nz = 0;
    for (int i = t; i < n; i += 256) {
      const double z = cur[i];
      dot = fma(x[i], z, dot);
      ss = fma(z, z, ss);
      nz |= (z != 0.0);
    }
    block_sum_w2(dot, ss, red);
    lambda = dot;
    sstot = ss;
    if (!__syncthreads_or(nz)) {
      lambda = 0.0;
      zero = true;
      break;
    }
    src = cur;
  }
  if (warm && !zero)
    for (int i = t; i < n; i += 256) J.vec[i] = src[i];
  if (t == 0) J.out[J.target] = sqrt(fabs(lambda));
}

// Any job: rows read from memory each round, vectors in dynamic LDS.
template <class T>
__global__ __launch_bounds__(256) void k_bpower_sp(const BPowSpJob<T>* __restrict__ jobs, int maxit, double thr,
                                                   int warm, int nmax) {
  extern __shared__ double lds[];  // x, u, w: nmax doubles each
  __shared__ double red[8];
  const BPowSpJob<T> J = jobs[blockIdx.x];
  if (J.m == 0) {
    if (threadIdx.x == 0) J.out[J.target] = 1.0;
    return;
  }
  bpower_sp_job<T, 0>(J, maxit, thr, warm, lds, lds + nmax, lds + 2 * nmax, red);
}

// Jobs of MC factors (or none) of edge n <= 512: rows in registers, vectors
// at fixed LDS offsets (so that every gather is one LDS read at a constant
// offset from the entry's index).
template <class T, int MC>
__global__ __launch_bounds__(256, 4) void k_bpower_sp_reg(const BPowSpJob<T>* __restrict__ jobs, int maxit, double thr,
                                                       int warm) {
  __shared__ double x[BSP_PAD + 2], u[BSP_PAD + 2], w[BSP_PAD + 2];
  __shared__ double red[8];
  const BPowSpJob<T> J = jobs[blockIdx.x];
  if (J.m == 0) {
    if (threadIdx.x == 0) J.out[J.target] = 1.0;
    return;
  }
  if (threadIdx.x == 0) x[BSP_PAD] = u[BSP_PAD] = w[BSP_PAD] = 0.0;
  bpower_sp_job<T, MC>(J, maxit, thr, warm, x, u, w, red);
}

// Batched sparse x dense products (BSpmmJob): a workgroup computes the
// columns [j0, j0 + nc) of one job's output.  The factor's entry lists (the
// rows of op(S) for S on the left, its columns for S on the right) are
// staged in LDS once per workgroup and serve all its columns; with S on the
// left each column of X is staged too and gathered from LDS, with S on the
// right the lanes read the combined columns of X directly (coalesced).  Each
// output entry is one fma chain over the entries in their stored order, FP32
// (T) accumulation like the dense GEMM it replaces.  Factors with more
// entries than the LDS room read them from global memory instead.
constexpr int SPMM_LDS_BYTES = 16 * 1024;
constexpr int SPMM_XCOLS = 2048;  // entries of X staged per column block (S on the left)
constexpr int SPMM_CB = 4;        // columns per block with S on the right
constexpr int SPMM_EB = 4;        // entries of a column loaded together (S on the right)

template <class T>
struct SpmmCtx {
  const int* s_p;
  const int* ix;  // entry indices and values: the LDS copies, or the factor's arrays
  const T* vv;
  T* s_x;
  const __attribute__((address_space(1))) T* X;
  const __attribute__((address_space(1))) T* Cin;
  __attribute__((address_space(1))) T* Cout;
  int n, j0, nc, side, beta_mode;
  T alpha, beta;
};

// IX_LDS: the entry lists are in LDS (the pointers of ctx point there), else
// in global memory; two copies of the loops, so that no load goes through a
// pointer that may be either (a flat access, which waits with the LDS reads)
template <class T, bool IX_LDS>
__device__ __forceinline__ void bspmm_body(const SpmmCtx<T>& C) {
  typedef const __attribute__((address_space(IX_LDS ? 3 : 1))) int* PI;
  typedef const __attribute__((address_space(IX_LDS ? 3 : 1))) T* PV;
  const PI ix = (PI)C.ix;
  const PV vv = (PV)C.vv;
  const int n = C.n, j0 = C.j0, nc = C.nc, t = threadIdx.x;
  // the epilogue of up to 8 outputs of one row: every Cin value loaded
  // before the first store (Cin may be Cout: each output still reads its own
  // old value first), so the loads are in flight together
  auto out8 = [&](int i, int jbase, int m, const T* acc) {
    T old[8];
    if (C.beta_mode) {
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (c < m) old[c] = C.Cin[(size_t)(jbase + c) * n + i];
    }
#pragma unroll
    for (int c = 0; c < 8; ++c)
      if (c < m) {
        T v = C.alpha * acc[c];
        if (C.beta_mode) v = fma(C.beta, old[c], v);
        C.Cout[(size_t)(jbase + c) * n + i] = v;
      }
  };
  // the same with the old values already loaded (issued with the operand
  // loads, so that their latency is not paid again in the epilogue)
  auto out_pre = [&](int i, int jbase, int m, const T* acc, const T* old) {
#pragma unroll
    for (int c = 0; c < 8; ++c)
      if (c < m) {
        T v = C.alpha * acc[c];
        if (C.beta_mode) v = fma(C.beta, old[c], v);
        C.Cout[(size_t)(jbase + c) * n + i] = v;
      }
  };
  if (C.side == 0) {
    // (op(S) X)(i, j) = sum over row i of op(S): v * X(k, j); cb columns of X
    // staged in LDS at a time, a row's entries read once for all of them
    const int cb = max(1, min(8, SPMM_XCOLS / n));
    // the old values of this thread's outputs in the block are loaded with the
    // staging when they fit 8 slots: one row (n <= 256, up to 8 columns) or
    // two rows of up to 4 columns (n <= 512 with blocks of 4 columns)
    const bool pre = C.beta_mode && (n <= 256 || (n <= 512 && cb <= 4));
    for (int jb = 0; jb < nc; jb += cb) {
      const int m = min(cb, nc - jb);
      T old[8];
      {
        // staging: all of this thread's loads first, then the LDS writes
        constexpr int PER = SPMM_XCOLS / 256;
        T st[PER];
        const int tot = m * n;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
          const int e = t + 256 * u;
          if (e < tot) st[u] = C.X[(size_t)(j0 + jb) * n + e];
        }
        if (pre) {
#pragma unroll
          for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              // slot r * 4 + c: row t + 256 r, column c (m <= 4 when there are two rows)
              const int i = t + 256 * r;
              if (i < n && c < m) old[r * 4 + c] = C.Cin[(size_t)(j0 + jb + c) * n + i];
            }
          if (n <= 256)
#pragma unroll
            for (int c = 4; c < 8; ++c)
              if (t < n && c < m) old[c] = C.Cin[(size_t)(j0 + jb + c) * n + t];
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
          const int e = t + 256 * u;
          if (e < tot) C.s_x[e] = st[u];
        }
      }
      __syncthreads();
      if (pre) {
        // at most two rows per thread here
        for (int r = 0; r < 2; ++r) {
          const int i = t + 256 * r;
          if (i >= n) break;
          T acc[8];
#pragma unroll
          for (int c = 0; c < 8; ++c) acc[c] = T(0);
          for (int q = C.s_p[i], qe = C.s_p[i + 1]; q < qe; ++q) {
            const int k = ix[q];
            const T v = vv[q];
#pragma unroll
            for (int c = 0; c < 8; ++c)
              if (c < m) acc[c] = fma(v, C.s_x[c * n + k], acc[c]);
          }
          out_pre(i, j0 + jb, m, acc, old + 4 * r);
        }
        __syncthreads();  // s_x is rewritten for the next block
        continue;
      }
      for (int i = t; i < n; i += 256) {
        T acc[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[c] = T(0);
        for (int q = C.s_p[i], qe = C.s_p[i + 1]; q < qe; ++q) {
          const int k = ix[q];
          const T v = vv[q];
#pragma unroll
          for (int c = 0; c < 8; ++c)
            if (c < m) acc[c] = fma(v, C.s_x[c * n + k], acc[c]);
        }
        out8(i, j0 + jb, m, acc);
      }
      __syncthreads();  // s_x is rewritten for the next block
    }
  } else {
    // (X op(S))(i, j) = sum over column j of op(S): X(i, k) * v; the X values of
    // SPMM_EB entries of each of SPMM_CB columns are loaded before their fmas
    for (int jb = 0; jb < nc; jb += SPMM_CB) {
      const int m = min(SPMM_CB, nc - jb);
      for (int i = t; i < n; i += 256) {
        T xv[SPMM_CB][SPMM_EB];
        int qa[SPMM_CB], qe[SPMM_CB];
#pragma unroll
        for (int c = 0; c < SPMM_CB; ++c) {
          const int j = min(j0 + jb + c, j0 + nc - 1);
          qa[c] = C.s_p[j];
          qe[c] = C.s_p[j + 1];
#pragma unroll
          for (int u = 0; u < SPMM_EB; ++u) {
            // slots past the column's entries read some valid row of X (the
            // index is clamped, also for a factor without entries) and are not used
            const int q = min(qa[c] + u, max(qe[c] - 1, 0));
            const int k = min(max(ix[q], 0), n - 1);
            xv[c][u] = C.X[(size_t)k * n + i];
          }
        }
        // the old values of the outputs, loaded with the X values
        T old[8];
        if (C.beta_mode) {
#pragma unroll
          for (int c = 0; c < SPMM_CB; ++c)
            if (c < m) old[c] = C.Cin[(size_t)(j0 + jb + c) * n + i];
        }
        T acc[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[c] = T(0);
#pragma unroll
        for (int c = 0; c < SPMM_CB; ++c) {
          if (c >= m) break;
#pragma unroll
          for (int u = 0; u < SPMM_EB; ++u)
            if (qa[c] + u < qe[c]) acc[c] = fma(xv[c][u], vv[qa[c] + u], acc[c]);
          // longer columns: the rest of the entries SPMM_EB at a time, loads first
          for (int q0 = qa[c] + SPMM_EB; q0 < qe[c]; q0 += SPMM_EB) {
            T xr[SPMM_EB];
#pragma unroll
            for (int u = 0; u < SPMM_EB; ++u) {
              const int q = min(q0 + u, qe[c] - 1);
              xr[u] = C.X[(size_t)ix[q] * n + i];
            }
#pragma unroll
            for (int u = 0; u < SPMM_EB; ++u)
              if (q0 + u < qe[c]) acc[c] = fma(xr[u], vv[q0 + u], acc[c]);
          }
        }
        out_pre(i, j0 + jb, m, acc, old);
      }
    }
  }
}

template <class T>
__global__ __launch_bounds__(256, 4) void k_bspmm(const BSpmmJob<T>* __restrict__ jobs, const int3* __restrict__ tiles,
                                                  BScal sc, int ntiles, int per) {
  constexpr int ENT = SPMM_LDS_BYTES / (sizeof(int) + sizeof(T));
  __shared__ int s_p[BATCH_NMAX + 1];
  __shared__ int s_ix[ENT];
  __shared__ T s_v[ENT];
  __shared__ T s_x[SPMM_XCOLS];
  const int tix = xcd_tile(per);
  if (tix >= ntiles) return;
  const int3 tl = tiles[tix];
  const BSpmmJob<T> J = jobs[tl.x];
  const int n = J.n, t = threadIdx.x;
  const bool rows_of_S = (J.side == 0) != (J.transS != 0);  // the entry lists are rows of S (CSR), else columns (CSC)
  const int* p0 = rows_of_S ? J.S.rp : J.S.cp;
  const int* ix = rows_of_S ? J.S.ci : J.S.ri;
  const T* vv = rows_of_S ? J.S.v : J.S.cv;
  const int nnz = p0[n];
  const bool lds = nnz <= ENT;
  for (int i = t; i <= n; i += 256) s_p[i] = p0[i];
  if (lds)
    for (int q = t; q < nnz; q += 256) {
      s_ix[q] = ix[q];
      s_v[q] = vv[q];
    }
  __syncthreads();
  SpmmCtx<T> C;
  C.s_p = s_p;
  C.ix = lds ? s_ix : ix;
  C.vv = lds ? s_v : vv;
  C.s_x = s_x;
  C.X = (const __attribute__((address_space(1))) T*)J.X;
  C.Cin = (const __attribute__((address_space(1))) T*)J.Cin;
  C.Cout = (__attribute__((address_space(1))) T*)J.Cout;
  C.n = n;
  C.j0 = tl.y;
  C.nc = tl.z;
  C.side = J.side;
  C.beta_mode = J.beta_mode;
  C.alpha = job_alpha<T>(J.alpha_mode, J.target, sc);
  C.beta = J.beta_mode == 0 ? T(0) : (J.beta_mode == 1 ? T(-1) : T(1));
  if (lds)
    bspmm_body<T, true>(C);
  else
    bspmm_body<T, false>(C);
}

// ---------------------------------------------------------------- projections
// One 1024-thread workgroup per target.  SP: keep the k largest magnitudes
// (radix select on the magnitude bits; among entries equal to the k-th
// largest the lowest indices are kept), then divide by the Frobenius norm.
// SUPP: multiply by the 0/1 mask, then normalise.

template <class T> struct KeyOf;
template <> struct KeyOf<float> {
  using K = unsigned;
  static constexpr int BITS = 32;
  __device__ static K key(float v) { return __float_as_uint(v) & 0x7fffffffu; }
};
template <> struct KeyOf<double> {
  using K = unsigned long long;
  static constexpr int BITS = 64;
  __device__ static K key(double v) { return (unsigned long long)__double_as_longlong(v) & 0x7fffffffffffffffull; }
};

constexpr int BSEL_BITS = 11;  // radix-select digit width (bins in LDS: 2048)

template <class T, int MODE>  // MODE 0: SP, 1: SUPP
__global__ __launch_bounds__(1024) void k_bprox(const BProxJob<T>* __restrict__ jobs) {
  using KO = KeyOf<T>;
  using K = typename KO::K;
  __shared__ unsigned hist[1 << BSEL_BITS];
  __shared__ unsigned wtot[16];
  __shared__ T red[1024];
  __shared__ unsigned scan[1024];
  __shared__ K s_prefix;
  __shared__ long long s_kk;
  __shared__ unsigned s_eq;
  // the kept entries of the untied path, so that the normalisation scales
  // them alone instead of every entry
  constexpr int KEEP_LIST = 8192;
  __shared__ int kept[KEEP_LIST];
  __shared__ int s_nkept;
  const BProxJob<T> J = jobs[blockIdx.x];
  const int t = threadIdx.x;
  const int64_t len = J.len;
  // global (not flat) accesses: their waits do not wait for the LDS atomics
  typedef __attribute__((address_space(1))) T* GT;
  GT M = (GT)J.M;
  // ss: this thread's sum of squares over i = t, t + 1024, ... of the
  // projected entries (fma in that order, as the separate pass below)
  T ss = T(0);
  bool have_ss = false;
  if (t == 0) s_nkept = -1;  // no list
  __syncthreads();
  if (MODE == 1) {
    for (int64_t i = t; i < len; i += 1024) M[i] = M[i] * J.mask[i];
  } else if (J.k < len) {
    if (J.k <= 0) {
      for (int64_t i = t; i < len; i += 1024) M[i] = T(0);
    } else {
      K prefix = 0, mask = 0;
      long long kk = J.k;  // rank still to find among keys matching prefix
      // digits of up to BSEL_BITS bits from the top: 3 passes for FP32 keys,
      // 6 for FP64
      for (int hi = KO::BITS; hi > 0; hi -= BSEL_BITS) {
        const int shift = hi > BSEL_BITS ? hi - BSEL_BITS : 0;
        const int nb = 1 << (hi - shift);
        const K dmask = (K)(nb - 1);
        for (int b = t; b < nb; b += 1024) hist[b] = 0;
        __syncthreads();
        // eight entries per thread loaded before they are binned, so that the
        // loads are in flight together
        int64_t i0 = t;
        for (; i0 + 7 * 1024 < len; i0 += 8 * 1024) {
          K kv[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) kv[u] = KO::key(M[i0 + u * 1024]);
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if ((kv[u] & mask) == prefix) atomicAdd(&hist[(kv[u] >> shift) & dmask], 1u);
        }
        for (int64_t i = i0; i < len; i += 1024) {
          const K key = KO::key(M[i]);
          if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & dmask], 1u);
        }
        __syncthreads();
        // the bin holding the kk-th largest key: thread t owns bins
        // [lo, lo + per); suffix sums over the threads (lane shuffles, then
        // the 16 wave totals) give each thread the count above its bins, and
        // the one thread whose bins hold rank kk walks them from the top
        const int per = nb >= 1024 ? nb / 1024 : 1;
        const int lo = t * per;
        unsigned own = 0;
        if (lo < nb)
          for (int q = 0; q < per; ++q) own += hist[lo + q];
        unsigned v = own;
        for (int off = 1; off < 64; off <<= 1) {
          const unsigned o = __shfl_down(v, off);
          if ((t & 63) + off < 64) v += o;
        }
        if ((t & 63) == 0) wtot[t >> 6] = v;
        if (t == 0) {
          // the bin search below always finds the bin (kk is at least 1 and at
          // most the keys matching prefix); these values are only a default
          s_kk = kk;
          s_prefix = prefix;
          s_eq = hist[0];
        }
        __syncthreads();
        unsigned above = 0;
        for (int w2 = (t >> 6) + 1; w2 < 16; ++w2) above += wtot[w2];
        const long long incl = (long long)v + above, excl = incl - own;
        if (lo < nb && excl < kk && kk <= incl) {
          long long cum = excl;
          int b = lo + per - 1;
          for (; b > lo; --b) {
            if (cum + hist[b] >= kk) break;
            cum += hist[b];
          }
          s_kk = kk - cum;
          s_prefix = prefix | ((K)b << shift);
          s_eq = hist[b];  // after the last digit: the entries equal to the k-th largest key
        }
        __syncthreads();
        kk = s_kk;
        prefix = s_prefix;
        mask |= (dmask << shift);
        __syncthreads();
      }
      // keys > prefix are kept; kk entries equal to prefix are kept, lowest index first
      if ((long long)s_eq == kk) {
        // every entry equal to the k-th largest key is kept: no ranking, and
        // the sum of squares is taken on the way
        if (t == 0) s_nkept = 0;
        __syncthreads();
        auto one = [&](int64_t i, T x) {
          const bool keep = KO::key(x) >= prefix;
          if (!keep) M[i] = T(0);
          const T y = keep ? x : T(0);
          ss = fma(y, y, ss);
          if (keep) {
            const int slot = atomicAdd(&s_nkept, 1);
            if (slot < KEEP_LIST) kept[slot] = (int)i;
          }
        };
        // eight loads in flight, then the entries in order (the order of the
        // sum of squares is unchanged)
        int64_t i0 = t;
        for (; i0 + 7 * 1024 < len; i0 += 8 * 1024) {
          T xv[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) xv[u] = M[i0 + u * 1024];
#pragma unroll
          for (int u = 0; u < 8; ++u) one(i0 + u * 1024, xv[u]);
        }
        for (int64_t i = i0; i < len; i += 1024) one(i, M[i]);
        have_ss = true;
      } else {
        const int64_t chunk = (len + 1023) / 1024;
        const int64_t c0 = t * chunk, c1 = min(len, c0 + chunk);
        unsigned eq = 0;
        for (int64_t i = c0; i < c1; ++i) eq += (KO::key(M[i]) == prefix);
        scan[t] = eq;
        __syncthreads();
        for (int s = 1; s < 1024; s <<= 1) {
          unsigned v = (t >= s) ? scan[t - s] : 0u;
          __syncthreads();
          scan[t] += v;
          __syncthreads();
        }
        long long before = (long long)scan[t] - eq;
        for (int64_t i = c0; i < c1; ++i) {
          const K key = KO::key(M[i]);
          bool keep = key > prefix;
          if (key == prefix) {
            keep = before < kk;
            ++before;
          }
          if (!keep) M[i] = T(0);
        }
      }
    }
  }
  __syncthreads();
  // Frobenius normalisation
  if (!have_ss)
    for (int64_t i = t; i < len; i += 1024) ss = fma(M[i], M[i], ss);
  red[t] = ss;
  __syncthreads();
  for (int s = 512; s >= 1; s >>= 1) {
    if (t < s) red[t] = red[t] + red[t + s];
    __syncthreads();
  }
  const T nrm = sqrt(red[0]);
  const int nk = s_nkept;
  const bool listed = nk >= 0 && nk <= KEEP_LIST;
  if (nrm > T(0)) {
    const T inv = T(1) / nrm;
    if (listed && isfinite(inv)) {
      // only the kept entries are nonzero (0 * inv = 0 for the others)
      for (int q = t; q < nk; q += 1024) M[kept[q]] = M[kept[q]] * inv;
    } else {
      for (int64_t i = t; i < len; i += 1024) M[i] = M[i] * inv;
    }
  }
  // the kept entries for the compressed form built next
  if (J.kcount) {
    const bool ok = listed && nk <= J.kcap;
    if (ok)
      for (int q = t; q < nk; q += 1024) J.klist[q] = kept[q];
    if (t == 0) *J.kcount = ok ? nk : -1;
  }
}

// ---------------------------------------------------------------- lambda / errors
// lambda_b = <A, S>_F / ||S||_F^2 (update_lambda, faust_palm4msa2020.hpp:442-493)

template <class T>
__global__ __launch_bounds__(1024) void k_blambda(const BLamJob<T>* __restrict__ jobs, double* __restrict__ lambda) {
  // 1024 threads, four independent partial sums per thread over a stride of
  // 4096 (loads of four iterations in flight), then wave shuffles and the 16
  // wave sums in order: a fixed order, the same result every run
  __shared__ double r1[16], r2[16];
  const BLamJob<T> J = jobs[blockIdx.x];
  const int t = threadIdx.x;
  typedef const __attribute__((address_space(1))) T* G;
  const G A = (G)J.A, S = (G)J.S;
  double tr[4] = {0.0, 0.0, 0.0, 0.0}, ss[4] = {0.0, 0.0, 0.0, 0.0};
  int64_t i = t;
  for (; i + 3 * 1024 < J.len; i += 4 * 1024) {
    T a[4], s[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = A[i + u * 1024];
      s[u] = S[i + u * 1024];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      tr[u] = fma((double)a[u], (double)s[u], tr[u]);
      ss[u] = fma((double)s[u], (double)s[u], ss[u]);
    }
  }
  for (; i < J.len; i += 1024) {
    const double sv = (double)S[i];
    tr[0] = fma((double)A[i], sv, tr[0]);
    ss[0] = fma(sv, sv, ss[0]);
  }
  double x = (tr[0] + tr[1]) + (tr[2] + tr[3]), y = (ss[0] + ss[1]) + (ss[2] + ss[3]);
  x = wave_sum_d(x);
  y = wave_sum_d(y);
  if ((t & 63) == 0) {
    r1[t >> 6] = x;
    r2[t >> 6] = y;
  }
  __syncthreads();
  if (t == 0) {
    double a1 = 0.0, a2 = 0.0;
    for (int w = 0; w < 16; ++w) {
      a1 += r1[w];
      a2 += r2[w];
    }
    if (a2 > 0.0) lambda[J.target] = a1 / a2;
  }
}

// rel_err_b = || lambda_b P - A ||_F / ||A||_F in FP64
template <class T>
__global__ __launch_bounds__(256) void k_brelerr(const BLamJob<T>* __restrict__ jobs, const double* __restrict__ lambda,
                                                 double* __restrict__ out) {
  __shared__ double r1[256], r2[256];
  const BLamJob<T> J = jobs[blockIdx.x];
  const int t = threadIdx.x;
  const double lam = lambda[J.target];
  double e = 0.0, a2 = 0.0;
  for (int64_t i = t; i < J.len; i += 256) {
    const double a = (double)J.A[i];
    const double d = lam * (double)J.S[i] - a;
    e = fma(d, d, e);
    a2 = fma(a, a, a2);
  }
  r1[t] = e;
  r2[t] = a2;
  __syncthreads();
  for (int s = 128; s >= 1; s >>= 1) {
    if (t < s) {
      r1[t] += r1[t + s];
      r2[t] += r2[t + s];
    }
    __syncthreads();
  }
  if (t == 0) out[J.target] = sqrt(r1[0]) / sqrt(r2[0]);
}

template <class T>
__global__ void k_bcopy(T* __restrict__ dst, const T* __restrict__ src, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void k_cast_f32_f64(double* __restrict__ dst, const float* __restrict__ src, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (double)src[i];
}
__global__ void k_mask_nz(double* __restrict__ mask, const float* __restrict__ src, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) mask[i] = src[i] != 0.0f ? 1.0 : 0.0;
}
void launch_cast_f32_f64(double* dst, const float* src, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cast_f32_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, src, n);
  FH_LAUNCH_CHECK();
}
void launch_mask_nonzero_f32(double* mask, const float* src, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mask_nz, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, mask, src, n);
  FH_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- host launchers
// tiles per XCD for xcd_tile, or 0 (tile = workgroup) under FH_BATCH_XCD=0
static int xcd_per(int ntiles) {
  static const bool on = [] {
    const char* e = std::getenv("FH_BATCH_XCD");
    return e == nullptr || std::string(e) != "0";
  }();
  return on ? (ntiles + 7) / 8 : 0;
}

template <class T>
void launch_bgemm(const BGemmJob<T>* jobs, const int3* tiles, int ntiles, int tile, const BScal& sc, hipStream_t st) {
  if (ntiles == 0) return;
  const int per = xcd_per(ntiles);
  const unsigned grid = per > 0 ? 8u * per : (unsigned)ntiles;
  if (tile == BGEMM_MFMA) {
    if constexpr (std::is_same<T, float>::value) {
      hipLaunchKernelGGL(k_bgemm_mfma_f32, dim3(grid), dim3(256), 0, st, jobs, tiles, sc, ntiles, per);
      FH_LAUNCH_CHECK();
      return;
    } else {
      throw Error(FH_ERR_ARG, "batched GEMM: the MFMA tile is FP32 only");
    }
  }
  if (tile == 32)
    hipLaunchKernelGGL((k_bgemm<T, 32>), dim3(grid), dim3(256), 0, st, jobs, tiles, sc, ntiles, per);
  else
    hipLaunchKernelGGL((k_bgemm<T, 64>), dim3(grid), dim3(256), 0, st, jobs, tiles, sc, ntiles, per);
  FH_LAUNCH_CHECK();
}

template <class T>
void launch_bpower(const BPowJob<T>* jobs, int njobs, int maxit, double thr, int warm, hipStream_t st) {
  if (njobs == 0) return;
  hipLaunchKernelGGL(k_bpower<T>, dim3(njobs), dim3(256), 0, st, jobs, maxit, thr, warm);
  FH_LAUNCH_CHECK();
}

template <class T>
void launch_bspmm(const BSpmmJob<T>* jobs, const int3* tiles, int ntiles, const BScal& sc, hipStream_t st) {
  if (ntiles == 0) return;
  const int per = xcd_per(ntiles);
  const unsigned grid = per > 0 ? 8u * per : (unsigned)ntiles;
  hipLaunchKernelGGL(k_bspmm<T>, dim3(grid), dim3(256), 0, st, jobs, tiles, sc, ntiles, per);
  FH_LAUNCH_CHECK();
}

template <class T>
void launch_bsparse(const BSpBuildJob<T>* jobs, int njobs, hipStream_t st) {
  if (njobs == 0) return;
  hipLaunchKernelGGL(k_bsparse<T>, dim3(njobs), dim3(256), 0, st, jobs);
  FH_LAUNCH_CHECK();
}

template <class T>
void launch_bpower_sp(const BPowSpJob<T>* jobs, int njobs, int nmax, int mc, int maxit, double thr, int warm,
                      hipStream_t st) {
  if (njobs == 0) return;
  if (nmax < 1 || nmax > BATCH_NMAX) throw Error(FH_ERR_ARG, "sparse power iteration: target edge out of range");
  if (mc != 0 && nmax > 512) throw Error(FH_ERR_ARG, "sparse power iteration: register rows need edges <= 512");
  if (mc == 1)
    hipLaunchKernelGGL((k_bpower_sp_reg<T, 1>), dim3(njobs), dim3(256), 0, st, jobs, maxit, thr, warm);
  else if (mc == 2)
    hipLaunchKernelGGL((k_bpower_sp_reg<T, 2>), dim3(njobs), dim3(256), 0, st, jobs, maxit, thr, warm);
  else
    hipLaunchKernelGGL(k_bpower_sp<T>, dim3(njobs), dim3(256), (size_t)3 * nmax * sizeof(double), st, jobs, maxit,
                       thr, warm, nmax);
  FH_LAUNCH_CHECK();
}

template <class T>
void launch_bprox(const BProxJob<T>* jobs, int njobs, int mode, hipStream_t st) {
  if (njobs == 0) return;
  if (mode == 0)
    hipLaunchKernelGGL((k_bprox<T, 0>), dim3(njobs), dim3(1024), 0, st, jobs);
  else
    hipLaunchKernelGGL((k_bprox<T, 1>), dim3(njobs), dim3(1024), 0, st, jobs);
  FH_LAUNCH_CHECK();
}

template <class T>
void launch_blambda(const BLamJob<T>* jobs, int njobs, double* lambda, hipStream_t st) {
  if (njobs == 0) return;
  hipLaunchKernelGGL(k_blambda<T>, dim3(njobs), dim3(1024), 0, st, jobs, lambda);
  FH_LAUNCH_CHECK();
}

template <class T>
void launch_brelerr(const BLamJob<T>* jobs, int njobs, const double* lambda, double* out, hipStream_t st) {
  if (njobs == 0) return;
  hipLaunchKernelGGL(k_brelerr<T>, dim3(njobs), dim3(256), 0, st, jobs, lambda, out);
  FH_LAUNCH_CHECK();
}

template <class S>
__global__ __launch_bounds__(256) void k_bcast(const BCastJob* __restrict__ jobs) {
  const BCastJob J = jobs[blockIdx.y];
  const S* src = static_cast<const S*>(J.src);
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < J.len; i += (int64_t)gridDim.x * 256)
    J.dst[i] = (double)src[i];
}

void launch_bcast(const BCastJob* jobs, int njobs, bool src_f32, hipStream_t st) {
  if (njobs == 0) return;
  const dim3 grid(64, (unsigned)njobs);
  if (src_f32)
    hipLaunchKernelGGL(k_bcast<float>, grid, dim3(256), 0, st, jobs);
  else
    hipLaunchKernelGGL(k_bcast<double>, grid, dim3(256), 0, st, jobs);
  FH_LAUNCH_CHECK();
}

template <class T>
void launch_bcopy(T* dst, const T* src, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_bcopy<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, src, n);
  FH_LAUNCH_CHECK();
}

#define FH_INST(T)                                                                                            \
  template void launch_bgemm<T>(const BGemmJob<T>*, const int3*, int, int, const BScal&, hipStream_t);       \
  template void launch_bpower<T>(const BPowJob<T>*, int, int, double, int, hipStream_t);                    \
  template void launch_bprox<T>(const BProxJob<T>*, int, int, hipStream_t);                                 \
  template void launch_bsparse<T>(const BSpBuildJob<T>*, int, hipStream_t);                                 \
  template void launch_bspmm<T>(const BSpmmJob<T>*, const int3*, int, const BScal&, hipStream_t);           \
  template void launch_bpower_sp<T>(const BPowSpJob<T>*, int, int, int, int, double, int, hipStream_t);    \
  template void launch_blambda<T>(const BLamJob<T>*, int, double*, hipStream_t);                            \
  template void launch_brelerr<T>(const BLamJob<T>*, int, const double*, double*, hipStream_t);             \
  template void launch_bcopy<T>(T*, const T*, int64_t, hipStream_t);
FH_INST(float)
FH_INST(double)

}  // namespace fh
