// k_prox.hip -- sp / spcol / splin / splincol / supp / normcol / normlin projectors (FP64).
//
// The selected support must equal the reference's std::partial_sort result
// bit for bit, ties included.  A workgroup radix select finds the k-th largest
// magnitude T of each line (row or column, staged in LDS) or of the whole
// matrix.  When the entries equal to T are all needed to reach k, the support
// is {|x| >= T} whatever the order of the scan; only when ties at T straddle
// the k boundary does the exact heap selection of heap_select.hpp run, on one
// lane, to pick the tied entries partial_sort keeps.  Zero magnitudes never
// need it: selecting a zero leaves a zero.  The selection is followed by a
// canonical Frobenius normalisation (faust_prox.hpp:72-200;
// MatDense::normalize, faust_MatDense.hpp:1226-1250).
#include <algorithm>

#include "fh_internal.hpp"
#include "heap_select.hpp"

namespace fh {

// Magnitude key: the bits of |x| order non-negative doubles like their values.
__device__ inline unsigned long long mag_key(double x) {
  return (unsigned long long)__double_as_longlong(fabs(x));
}

// Workgroup radix select (8-bit digits, most significant first) of the k-th
// largest of len keys, 1 <= k <= len.  On return *T is a threshold such that
// the keys >= *T are exactly the k largest, unless the function returns true:
// then *T is the k-th largest key itself and the keys equal to it straddle
// the k boundary, so which of them partial_sort keeps depends on positions
// (the caller runs the exact heap selection); a zero T is never reported as
// ambiguous, since keeping a zero or not gives the same output.  hist: 256 LDS counters; sh: 3
// LDS ints.  Each lane caches two (digit, count) pairs before its LDS atomics,
// which keeps the passes where most keys share a digit (the exponent bits)
// free of atomic contention.
template <class KeyAt>
__device__ bool wg_select_threshold(KeyAt key_at, int len, int k, unsigned* hist, int* sh,
                                    unsigned long long* T) {
  const int tid = threadIdx.x, nt = blockDim.x;
  unsigned long long prefix = 0, mask = 0;
  int kk = k;  // rank of the target among the candidates, from the top
  for (int shift = 56; shift >= 0; shift -= 8) {
    for (int b = tid; b < 256; b += nt) hist[b] = 0u;
    __syncthreads();
    unsigned d0 = 256, c0 = 0, d1 = 256, c1 = 0;
    for (int i = tid; i < len; i += nt) {
      const unsigned long long x = key_at(i);
      if ((x & mask) != prefix) continue;
      const unsigned d = (unsigned)(x >> shift) & 255u;
      if (d == d0) {
        ++c0;
      } else if (d == d1) {
        ++c1;
      } else {
        if (c1) atomicAdd(&hist[d1], c1);
        d1 = d0;
        c1 = c0;
        d0 = d;
        c0 = 1;
      }
    }
    if (c0) atomicAdd(&hist[d0], c0);
    if (c1) atomicAdd(&hist[d1], c1);
    __syncthreads();
    if (tid < 64) {
      // lane l owns bins 255-4l .. 252-4l; inclusive scan of their sums from the top
      const int hi = 255 - 4 * tid;
      const int s = (int)(hist[hi] + hist[hi - 1] + hist[hi - 2] + hist[hi - 3]);
      int incl = s;
      for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(incl, d, 64);
        if (tid >= d) incl += o;
      }
      const unsigned long long bal = __ballot(incl >= kk);
      const int L = __ffsll((unsigned long long)bal) - 1;
      if (tid == L) {
        int acc = incl - s, b = hi, c = 0;
        for (int j = 0; j < 4; ++j, --b) {
          c = (int)hist[b];
          if (acc + c >= kk) break;
          acc += c;
        }
        sh[0] = b;
        sh[1] = kk - acc;
        sh[2] = c;
      }
    }
    __syncthreads();
    const int b = sh[0], cnt = sh[2];
    kk = sh[1];
    prefix |= (unsigned long long)b << shift;
    mask |= 255ull << shift;
    if (cnt == kk) {  // every candidate left in the bin is taken
      *T = prefix;
      return false;
    }
  }
  *T = prefix;
  return prefix != 0;  // tied zeros: every choice gives the same output
}

// One 64-lane workgroup per line.  along_rows: line i is row i (splin);
// otherwise line j is column j (spcol).  Dynamic LDS holds the line values and
// the heap when they fit (lds_ok), else per-line global scratch is used.  The
// radix threshold decides the support unless ties straddle the k boundary;
// only those lines run the heap selection on one lane.
__global__ __launch_bounds__(64) void k_prox_lines(const double* __restrict__ M, int m, int n, int along_rows, int k,
                                                   double* __restrict__ out, int lds_ok, double* __restrict__ gvals,
                                                   int* __restrict__ gidx, double* __restrict__ gabs) {
  extern __shared__ double smem[];
  __shared__ unsigned hist[256];
  __shared__ int sh[3];
  const int line = blockIdx.x;
  const int len = along_rows ? n : m;
  double* vals;
  double* habs;
  int* hidx;
  if (lds_ok) {
    vals = smem;
    habs = smem + len;
    hidx = reinterpret_cast<int*>(smem + len + k);
  } else {
    vals = gvals + (size_t)line * len;
    habs = gabs + (size_t)line * k;
    hidx = gidx + (size_t)line * k;
  }
  const int lane = threadIdx.x;
  for (int t = lane; t < len; t += 64) vals[t] = along_rows ? M[(size_t)t * m + line] : M[(size_t)line * m + t];
  __syncthreads();
  unsigned long long T;
  const bool ambiguous = wg_select_threshold([vals](int i) { return mag_key(vals[i]); }, len, k, hist, sh, &T);
  for (int t = lane; t < len; t += 64) {
    const double x = (!ambiguous && mag_key(vals[t]) >= T) ? vals[t] : 0.0;
    if (along_rows)
      out[(size_t)t * m + line] = x;
    else
      out[(size_t)line * m + t] = x;
  }
  if (!ambiguous) return;
  __syncthreads();
  if (lane == 0) heap_select<double>(len, k, [vals](int i) { return fabs(vals[i]); }, hidx, habs);
  __syncthreads();
  for (int q = lane; q < k; q += 64) {
    int t = hidx[q];
    if (along_rows)
      out[(size_t)t * m + line] = vals[t];
    else
      out[(size_t)line * m + t] = vals[t];
  }
}

// Global top-k over all m*n entries in column-major order (prox_sp).
struct SpSel {
  unsigned long long T;
  int ambiguous;
};

__global__ __launch_bounds__(1024) void k_sp_threshold(const double* __restrict__ M, int N, int k,
                                                       SpSel* __restrict__ sel) {
  __shared__ unsigned hist[256];
  __shared__ int sh[3];
  unsigned long long T;
  const bool amb = wg_select_threshold([M](int i) { return mag_key(M[i]); }, N, k, hist, sh, &T);
  if (threadIdx.x == 0) {
    sel->T = T;
    sel->ambiguous = amb ? 1 : 0;
  }
}

// exact heap selection on one lane, only when the threshold left ties open
__global__ __launch_bounds__(64) void k_prox_global_select(const double* __restrict__ M, int64_t N, int k,
                                                           const SpSel* __restrict__ sel, int* __restrict__ hidx,
                                                           double* __restrict__ habs) {
  if (!sel->ambiguous) return;
  if (threadIdx.x == 0) heap_select<double>((int)N, k, [M](int i) { return fabs(M[i]); }, hidx, habs);
}

__global__ void k_sp_apply(const double* __restrict__ M, int64_t N, const SpSel* __restrict__ sel,
                           double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double x = M[i];
  out[i] = (!sel->ambiguous && mag_key(x) >= sel->T) ? x : 0.0;
}

__global__ void k_scatter_sel(const double* __restrict__ M, const int* __restrict__ idx, int k,
                              const SpSel* __restrict__ sel, double* __restrict__ out) {
  if (!sel->ambiguous) return;
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < k) out[idx[q]] = M[idx[q]];
}

__global__ void k_pos(double* __restrict__ M, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (float)M[i] < 0.0f) M[i] = 0.0;
}

__global__ void k_union(const double* __restrict__ Mc, const double* __restrict__ Ml, double* __restrict__ out,
                        int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double c = Mc[i];
    double l = (c != 0.0) ? 0.0 : Ml[i];
    out[i] = c + l;
  }
}

__global__ void k_fill_zero(double* __restrict__ M, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) M[i] = 0.0;
}

__global__ void k_hadamard(double* __restrict__ M, const double* __restrict__ S, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) M[i] = M[i] * S[i];
}

// normcol / normlin (faust_prox.hpp:205-255): one 256-lane workgroup per
// line scales it to 2-norm s.  The line's squared norm is the canonical dot
// (4096-element chunks, lane t takes t, t+256, ..., a halving tree per chunk,
// chunk partials added in order), so the scale s/||line|| has the oracle's bits.
__global__ __launch_bounds__(256) void k_line_norm_scale(double* __restrict__ M, int m, int n, int along_rows,
                                                         double s) {
  __shared__ double p[256];
  const int line = blockIdx.x, t = threadIdx.x;
  const int len = along_rows ? n : m;
  const int64_t stride = along_rows ? m : 1;
  double* x = along_rows ? M + line : M + (int64_t)line * m;
  double tot = 0.0;
  for (int c0 = 0; c0 < len; c0 += RED_CHUNK) {
    double acc = 0.0;
    for (int j = 0; j < RED_CHUNK / RED_LANES; ++j) {
      const int i = c0 + j * RED_LANES + t;
      if (i < len) {
        const double v = x[i * stride];
        acc = fma(v, v, acc);
      }
    }
    p[t] = acc;
    __syncthreads();
    for (int h = RED_LANES / 2; h >= 1; h >>= 1) {
      if (t < h) p[t] = p[t] + p[t + h];
      __syncthreads();
    }
    tot = tot + p[0];
    __syncthreads();
  }
  double scale = sqrt(tot);
  if (scale != 0.0) scale = s / scale;
  for (int i = t; i < len; i += RED_LANES) x[i * stride] = x[i * stride] * scale;
}

static inline dim3 g1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

// top-k per line into `out` (dense, same shape as M)
static void prox_lines(const double* M, int m, int n, bool along_rows, int k, double* out, hipStream_t st) {
  const int nlines = along_rows ? m : n;
  const int len = along_rows ? n : m;
  if (nlines == 0) return;
  size_t lds = sizeof(double) * ((size_t)len + k) + sizeof(int) * (size_t)k;
  const size_t LDS_MAX = 60 * 1024;
  if (lds <= LDS_MAX) {
    hipLaunchKernelGGL(k_prox_lines, dim3(nlines), dim3(64), lds, st, M, m, n, along_rows ? 1 : 0, k, out, 1,
                       (double*)nullptr, (int*)nullptr, (double*)nullptr);
  } else {
    DBuf gv(sizeof(double) * (size_t)nlines * len), gi(sizeof(int) * (size_t)nlines * k),
        ga(sizeof(double) * (size_t)nlines * k);
    hipLaunchKernelGGL(k_prox_lines, dim3(nlines), dim3(64), 0, st, M, m, n, along_rows ? 1 : 0, k, out, 0,
                       gv.as<double>(), gi.as<int>(), ga.as<double>());
    FH_LAUNCH_CHECK();
    return;
  }
  FH_LAUNCH_CHECK();
}

static void normalize_inplace(double* M, int64_t N, hipStream_t st) {
  DBuf s(sizeof(double));
  reduce_dot_f64(M, M, N, s.as<double>(), st);
  double h = 0;
  FH_HIP(hipMemcpyAsync(&h, s.p, sizeof(double), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  if (h == 0.0) throw Error(FH_ERR_NUMERIC, "the norm is zero, can't normalize.");
  scale_f64(M, N, s.as<double>(), 0.0, 1, st);
}

void prox_f64(int name, double* M, int m, int n, int64_t k, bool normalized, bool pos, const double* supp,
              hipStream_t st) {
  const int64_t N = (int64_t)m * n;
  if (N == 0) return;
  auto zero_all = [&]() { FH_HIP(hipMemsetAsync(M, 0, sizeof(double) * N, st)); };
  switch (name) {
    case FH_PROX_SP: {
      if (pos) hipLaunchKernelGGL(k_pos, g1(N), dim3(256), 0, st, M, N);
      if (k <= 0)
        zero_all();
      else if (k < N) {
        FH_CHECK(N < (int64_t)1 << 31, "prox_sp: matrix too large");
        DBuf hi(sizeof(int) * k), ha(sizeof(double) * k), out(sizeof(double) * N), sel(sizeof(SpSel));
        hipLaunchKernelGGL(k_sp_threshold, dim3(1), dim3(1024), 0, st, M, (int)N, (int)k, sel.as<SpSel>());
        FH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_prox_global_select, dim3(1), dim3(64), 0, st, M, N, (int)k, sel.as<SpSel>(),
                           hi.as<int>(), ha.as<double>());
        FH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sp_apply, g1(N), dim3(256), 0, st, M, N, sel.as<SpSel>(), out.as<double>());
        FH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_scatter_sel, g1(k), dim3(256), 0, st, M, hi.as<int>(), (int)k, sel.as<SpSel>(),
                           out.as<double>());
        FH_LAUNCH_CHECK();
        FH_HIP(hipMemcpyAsync(M, out.p, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
      }
      break;
    }
    case FH_PROX_SPCOL:
    case FH_PROX_SPLIN: {
      const bool rows = (name == FH_PROX_SPLIN);
      if (pos) hipLaunchKernelGGL(k_pos, g1(N), dim3(256), 0, st, M, N);
      if (k <= 0)
        zero_all();
      else if (k < (rows ? n : m)) {
        DBuf out(sizeof(double) * N);
        prox_lines(M, m, n, rows, (int)k, out.as<double>(), st);
        FH_HIP(hipMemcpyAsync(M, out.p, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
      }
      break;
    }
    case FH_PROX_SPLINCOL: {
      // faust_prox.hpp:180-200: both selections run on the input (pos only
      // affects the matrix that is then overwritten)
      DBuf mc(sizeof(double) * N), ml(sizeof(double) * N);
      FH_HIP(hipMemcpyAsync(mc.p, M, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
      FH_HIP(hipMemcpyAsync(ml.p, M, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
      if (k <= 0) {
        FH_HIP(hipMemsetAsync(mc.p, 0, sizeof(double) * N, st));
        FH_HIP(hipMemsetAsync(ml.p, 0, sizeof(double) * N, st));
      } else {
        if (k < m) prox_lines(M, m, n, false, (int)k, mc.as<double>(), st);
        if (k < n) prox_lines(M, m, n, true, (int)k, ml.as<double>(), st);
      }
      hipLaunchKernelGGL(k_union, g1(N), dim3(256), 0, st, mc.as<double>(), ml.as<double>(), M, N);
      FH_LAUNCH_CHECK();
      break;
    }
    case FH_PROX_SUPP: {
      FH_CHECK(supp != nullptr, "prox_supp: missing support matrix");
      if (pos) hipLaunchKernelGGL(k_pos, g1(N), dim3(256), 0, st, M, N);
      hipLaunchKernelGGL(k_hadamard, g1(N), dim3(256), 0, st, M, supp, N);
      FH_LAUNCH_CHECK();
      break;
    }
    default:
      throw Error(FH_ERR_ARG, "prox: unknown constraint");
  }
  if (normalized) normalize_inplace(M, N, st);
}

void prox_real_f64(int name, double* M, int m, int n, double s, bool normalized, bool pos, hipStream_t st) {
  const int64_t N = (int64_t)m * n;
  FH_CHECK(name == FH_PROX_NORMCOL || name == FH_PROX_NORMLIN, "prox: unknown real-parameter constraint");
  if (N == 0) return;
  if (pos) {
    hipLaunchKernelGGL(k_pos, g1(N), dim3(256), 0, st, M, N);
    FH_LAUNCH_CHECK();
  }
  if (s < 0) throw Error(FH_ERR_ARG, "Faust::prox_normcol : s < 0 ");
  if (s == 0) {
    hipLaunchKernelGGL(k_fill_zero, g1(N), dim3(256), 0, st, M, N);
  } else {
    const bool rows = name == FH_PROX_NORMLIN;
    hipLaunchKernelGGL(k_line_norm_scale, dim3(rows ? m : n), dim3(256), 0, st, M, m, n, rows ? 1 : 0, s);
  }
  FH_LAUNCH_CHECK();
  if (normalized) normalize_inplace(M, N, st);
}

}  // namespace fh
