// k_prox.hip -- sp / spcol / splin / splincol / supp projectors (FP64).
//
// The selected support must equal the reference's std::partial_sort result
// bit for bit, ties included, so every line (row or column) runs the exact
// heap selection of heap_select.hpp on one lane, with the line staged in LDS
// by its wave.  The selection is followed by a canonical Frobenius
// normalisation (faust_prox.hpp:72-200; MatDense::normalize,
// faust_MatDense.hpp:1226-1250).
#include <algorithm>

#include "fh_internal.hpp"
#include "heap_select.hpp"

namespace fh {

// One 64-lane workgroup per line.  along_rows: line i is row i (splin);
// otherwise line j is column j (spcol).  Dynamic LDS holds the line values and
// the heap when they fit (lds_ok), else per-line global scratch is used.
__global__ __launch_bounds__(64) void k_prox_lines(const double* __restrict__ M, int m, int n, int along_rows, int k,
                                                   double* __restrict__ out, int lds_ok, double* __restrict__ gvals,
                                                   int* __restrict__ gidx, double* __restrict__ gabs) {
  extern __shared__ double smem[];
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
  for (int t = lane; t < len; t += 64) {
    double x = along_rows ? M[(size_t)t * m + line] : M[(size_t)line * m + t];
    vals[t] = x;
    if (along_rows)
      out[(size_t)t * m + line] = 0.0;
    else
      out[(size_t)line * m + t] = 0.0;
  }
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

// Global top-k over all m*n entries in column-major order (prox_sp): one lane.
__global__ __launch_bounds__(64) void k_prox_global_select(const double* __restrict__ M, int64_t N, int k,
                                                           int* __restrict__ hidx, double* __restrict__ habs) {
  if (threadIdx.x == 0) heap_select<double>((int)N, k, [M](int i) { return fabs(M[i]); }, hidx, habs);
}

__global__ void k_scatter_sel(const double* __restrict__ M, const int* __restrict__ idx, int k,
                              double* __restrict__ out) {
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

__global__ void k_hadamard(double* __restrict__ M, const double* __restrict__ S, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) M[i] = M[i] * S[i];
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
    FH_HIP(hipStreamSynchronize(st));  // scratch is returned to the pool below
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
  FH_HIP(hipStreamSynchronize(st));
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
        DBuf hi(sizeof(int) * k), ha(sizeof(double) * k), out(sizeof(double) * N);
        hipLaunchKernelGGL(k_prox_global_select, dim3(1), dim3(64), 0, st, M, N, (int)k, hi.as<int>(),
                           ha.as<double>());
        FH_LAUNCH_CHECK();
        FH_HIP(hipMemsetAsync(out.p, 0, sizeof(double) * N, st));
        hipLaunchKernelGGL(k_scatter_sel, g1(k), dim3(256), 0, st, M, hi.as<int>(), (int)k, out.as<double>());
        FH_LAUNCH_CHECK();
        FH_HIP(hipMemcpyAsync(M, out.p, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
        FH_HIP(hipStreamSynchronize(st));
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
        FH_HIP(hipStreamSynchronize(st));
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
      FH_HIP(hipStreamSynchronize(st));
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

}  // namespace fh
