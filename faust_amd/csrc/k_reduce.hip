// k_reduce.hip -- canonical FP64 reductions, elementwise kernels and
// dense <-> CSR/CSC conversions.
#include <algorithm>
#include <numeric>

#include "fh_internal.hpp"

namespace fh {

// ---------------------------------------------------------------- reductions
// One workgroup of 256 lanes per 4096-element chunk; lane t accumulates
// elements c*4096 + t + 256*m (m = 0..15), then a halving tree.
template <int MODE>  // 0: sum x, 1: dot x.y
__global__ __launch_bounds__(256) void k_reduce_chunks(const double* __restrict__ x, const double* __restrict__ y,
                                                       int64_t n, double* __restrict__ part) {
  __shared__ double p[RED_LANES];
  const int t = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * RED_CHUNK;
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < RED_CHUNK / RED_LANES; ++m) {
    int64_t i = base + (int64_t)m * RED_LANES + t;
    if (i < n) acc = MODE ? fma(x[i], y[i], acc) : acc + x[i];
  }
  p[t] = acc;
  __syncthreads();
  for (int s = RED_LANES / 2; s >= 1; s >>= 1) {
    if (t < s) p[t] = p[t] + p[t + s];
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = p[0];
}

__global__ void k_reduce_final(const double* __restrict__ part, int64_t nchunks, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double tot = 0.0;
    for (int64_t c = 0; c < nchunks; ++c) tot = tot + part[c];
    *out = tot;
  }
}

static void reduce_launch(int mode, const double* x, const double* y, int64_t n, double* out, hipStream_t st) {
  if (n <= 0) {
    FH_HIP(hipMemsetAsync(out, 0, sizeof(double), st));
    return;
  }
  int64_t nch = (n + RED_CHUNK - 1) / RED_CHUNK;
  DBuf part(sizeof(double) * nch);
  if (mode)
    hipLaunchKernelGGL(k_reduce_chunks<1>, dim3((unsigned)nch), dim3(RED_LANES), 0, st, x, y, n, part.as<double>());
  else
    hipLaunchKernelGGL(k_reduce_chunks<0>, dim3((unsigned)nch), dim3(RED_LANES), 0, st, x, y, n, part.as<double>());
  FH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, st, part.as<double>(), nch, out);
  FH_LAUNCH_CHECK();
}

void reduce_dot_f64(const double* x, const double* y, int64_t n, double* out, hipStream_t st) {
  reduce_launch(1, x, y, n, out, st);
}
void reduce_sum_f64(const double* x, int64_t n, double* out, hipStream_t st) { reduce_launch(0, x, nullptr, n, out, st); }

// diag[i] = sum_k A[k,i]*S[k,i] in ascending k (the diagonal of A^T S).
__global__ void k_diag_dot(const double* __restrict__ A, const double* __restrict__ S, int m, int n,
                           double* __restrict__ diag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* a = A + (size_t)i * m;
  const double* s = S + (size_t)i * m;
  double acc = 0.0;
  for (int k = 0; k < m; ++k) acc = fma(a[k], s[k], acc);
  diag[i] = acc;
}

void diag_dot_f64(const double* A, const double* S, int m, int n, double* diag, hipStream_t st) {
  hipLaunchKernelGGL(k_diag_dot, dim3((n + 255) / 256), dim3(256), 0, st, A, S, m, n, diag);
  FH_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- elementwise
__global__ void k_axpby(double* __restrict__ y, const double* __restrict__ x, int64_t n, double a, double b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double t = a * x[i];
    y[i] = (b == 0.0) ? t : fma(b, y[i], t);
  }
}
void axpby_f64(double* y, const double* x, int64_t n, double a, double b, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpby, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, y, x, n, a, b);
  FH_LAUNCH_CHECK();
}

// mode 0: x *= hscale; mode 1: x *= 1/sqrt(*dscale) (Frobenius normalisation
// from a device sum of squares: nrm = sqrt(s), inv = 1.0/nrm, x *= inv)
__global__ void k_scale(double* __restrict__ x, int64_t n, const double* __restrict__ ds, double hs, int mode) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = hs;
  if (mode == 1) {
    double nrm = sqrt(*ds);
    s = 1.0 / nrm;
  }
  x[i] = x[i] * s;
}
void scale_f64(double* x, int64_t n, const double* dscale, double hscale, int mode, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, n, dscale, hscale, mode);
  FH_LAUNCH_CHECK();
}

__global__ void k_sub(double* __restrict__ y, const double* __restrict__ x, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = y[i] - x[i];
}
void sub_f64(double* y, const double* x, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sub, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, y, x, n);
  FH_LAUNCH_CHECK();
}

__global__ void k_fill(double* __restrict__ x, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}
void fill_f64(double* x, int64_t n, double v, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, n, v);
  FH_LAUNCH_CHECK();
}

__global__ void k_eye(double* __restrict__ x, int m, int n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (int64_t)m * n) {
    int r = (int)(i % m), c = (int)(i / m);
    x[i] = (r == c) ? 1.0 : 0.0;
  }
}
void eye_f64(double* x, int m, int n, hipStream_t st) {
  int64_t N = (int64_t)m * n;
  if (N <= 0) return;
  hipLaunchKernelGGL(k_eye, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, x, m, n);
  FH_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- conversions
// Dense -> CSR and CSC in chunks of 32 columns (rows for CSC) so that every
// thread scans 32 entries: counts per (row, chunk), a host prefix sum, then
// an independent fill per (row, chunk) in ascending column order.
constexpr int CV_CH = 32;

__global__ void k_count_row_chunks(const double* __restrict__ D, int m, int n, int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (i >= m) return;
  const int j1 = min(n, (c + 1) * CV_CH);
  int k = 0;
  for (int j = c * CV_CH; j < j1; ++j) k += (D[(size_t)j * m + i] != 0.0);
  cnt[(size_t)c * m + i] = k;
}
__global__ void k_count_col_chunks(const double* __restrict__ D, int m, int n, int* __restrict__ cnt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (j >= n) return;
  const double* col = D + (size_t)j * m;
  const int i1 = min(m, (c + 1) * CV_CH);
  int k = 0;
  for (int i = c * CV_CH; i < i1; ++i) k += (col[i] != 0.0);
  cnt[(size_t)c * n + j] = k;
}
__global__ void k_fill_row_chunks(const double* __restrict__ D, int m, int n, const int* __restrict__ off,
                                  int* __restrict__ ci, double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (i >= m) return;
  int p = off[(size_t)c * m + i];
  const int j1 = min(n, (c + 1) * CV_CH);
  for (int j = c * CV_CH; j < j1; ++j) {
    const double x = D[(size_t)j * m + i];
    if (x != 0.0) {
      ci[p] = j;
      v[p] = x;
      ++p;
    }
  }
}
__global__ void k_fill_col_chunks(const double* __restrict__ D, int m, int n, const int* __restrict__ off,
                                  int* __restrict__ ri, double* __restrict__ v) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (j >= n) return;
  const double* col = D + (size_t)j * m;
  int p = off[(size_t)c * n + j];
  const int i1 = min(m, (c + 1) * CV_CH);
  for (int i = c * CV_CH; i < i1; ++i) {
    const double x = col[i];
    if (x != 0.0) {
      ri[p] = i;
      v[p] = x;
      ++p;
    }
  }
}

// counts [nch][len] -> per-(chunk, line) write offsets and the line pointer array
static void chunk_offsets(const std::vector<int>& cnt, int nch, int len, std::vector<int>& off,
                          std::vector<int>& ptr) {
  ptr.assign(len + 1, 0);
  for (int c = 0; c < nch; ++c)
    for (int i = 0; i < len; ++i) ptr[i + 1] += cnt[(size_t)c * len + i];
  for (int i = 0; i < len; ++i) ptr[i + 1] += ptr[i];
  off.resize((size_t)nch * len);
  for (int i = 0; i < len; ++i) {
    int p = ptr[i];
    for (int c = 0; c < nch; ++c) {
      off[(size_t)c * len + i] = p;
      p += cnt[(size_t)c * len + i];
    }
  }
}

void dense_to_csr_f64(const double* D, int m, int n, DMat& out, hipStream_t st) {
  out.kind = K_CSR;
  out.dtype = FH_F64;
  out.m = m;
  out.n = n;
  const int nchr = std::max(1, (n + CV_CH - 1) / CV_CH);  // column chunks (CSR)
  const int nchc = std::max(1, (m + CV_CH - 1) / CV_CH);  // row chunks (CSC)
  std::vector<int> hr((size_t)nchr * m, 0), hc((size_t)nchc * n, 0);
  DBuf cr(sizeof(int) * std::max<size_t>(hr.size(), 1)), cc(sizeof(int) * std::max<size_t>(hc.size(), 1));
  if (m > 0 && n > 0) {
    hipLaunchKernelGGL(k_count_row_chunks, dim3((m + 255) / 256, nchr), dim3(256), 0, st, D, m, n, cr.as<int>());
    FH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_count_col_chunks, dim3((n + 255) / 256, nchc), dim3(256), 0, st, D, m, n, cc.as<int>());
    FH_LAUNCH_CHECK();
    FH_HIP(hipMemcpyAsync(hr.data(), cr.p, sizeof(int) * hr.size(), hipMemcpyDeviceToHost, st));
    FH_HIP(hipMemcpyAsync(hc.data(), cc.p, sizeof(int) * hc.size(), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
  }
  std::vector<int> offr, rp, offc, cp;
  chunk_offsets(hr, nchr, m, offr, rp);
  chunk_offsets(hc, nchc, n, offc, cp);
  const int64_t nnz = rp[m];
  out.nnz = nnz;
  out.rp.alloc(sizeof(int) * (m + 1));
  out.cp.alloc(sizeof(int) * (n + 1));
  FH_HIP(hipMemcpyAsync(out.rp.p, rp.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice, st));
  FH_HIP(hipMemcpyAsync(out.cp.p, cp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, st));
  out.ci.alloc(sizeof(int) * std::max<int64_t>(nnz, 1));
  out.v.alloc(sizeof(double) * std::max<int64_t>(nnz, 1));
  out.ri.alloc(sizeof(int) * std::max<int64_t>(nnz, 1));
  out.cv.alloc(sizeof(double) * std::max<int64_t>(nnz, 1));
  if (nnz > 0) {
    FH_HIP(hipMemcpyAsync(cr.p, offr.data(), sizeof(int) * offr.size(), hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(cc.p, offc.data(), sizeof(int) * offc.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_fill_row_chunks, dim3((m + 255) / 256, nchr), dim3(256), 0, st, D, m, n, cr.as<int>(),
                       out.ci.as<int>(), out.v.as<double>());
    FH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_fill_col_chunks, dim3((n + 255) / 256, nchc), dim3(256), 0, st, D, m, n, cc.as<int>(),
                       out.ri.as<int>(), out.cv.as<double>());
    FH_LAUNCH_CHECK();
  }
  out.has_csc = true;
  // the host vectors go out of scope: make sure the uploads have completed
  FH_HIP(hipStreamSynchronize(st));
}

__global__ void k_csr_scatter(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                              const double* __restrict__ v, double* __restrict__ D) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  for (int p = rp[i]; p < rp[i + 1]; ++p) D[(size_t)ci[p] * m + i] = v[p];
}

void csr_to_dense_f64(const DMat& S, double* D, hipStream_t st) {
  FH_HIP(hipMemsetAsync(D, 0, sizeof(double) * (size_t)S.m * S.n, st));
  if (S.nnz > 0) {
    hipLaunchKernelGGL(k_csr_scatter, dim3((S.m + 255) / 256), dim3(256), 0, st, S.m, S.rp.as<int>(),
                       S.ci.as<int>(), S.v.as<double>(), D);
    FH_LAUNCH_CHECK();
  }
}

__device__ __forceinline__ bool nz_val(double x) { return x != 0.0; }
__device__ __forceinline__ bool nz_val(float x) { return x != 0.0f; }
__device__ __forceinline__ bool nz_val(uint16_t x) { return (x & 0x7fffu) != 0; }  // BF16 bits: +-0 are zero

// grid-stride count: per-thread counts, a block sum in LDS, one atomic per
// workgroup (an atomic per wave on one counter serialised ~16K atomics for a
// 1024^2 matrix)
template <class T>
__global__ __launch_bounds__(256) void k_count_nnz(const T* __restrict__ D, int64_t n,
                                                   unsigned long long* __restrict__ out) {
  __shared__ unsigned int part[256];
  unsigned int c = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    c += nz_val(D[i]) ? 1u : 0u;
  part[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && part[0]) atomicAdd(out, (unsigned long long)part[0]);
}

int64_t count_nnz_f64(const double* D, int64_t n, hipStream_t st) { return count_nnz(D, FH_F64, n, st); }

int64_t count_nnz(const void* D, int dtype, int64_t n, hipStream_t st) {
  if (n <= 0) return 0;
  DBuf c(sizeof(unsigned long long));
  FH_HIP(hipMemsetAsync(c.p, 0, sizeof(unsigned long long), st));
  const dim3 grid((unsigned)std::min<int64_t>(1024, (n + 255) / 256));
  if (dtype == FH_F64)
    hipLaunchKernelGGL(k_count_nnz<double>, grid, dim3(256), 0, st, (const double*)D, n, c.as<unsigned long long>());
  else if (dtype == FH_F32)
    hipLaunchKernelGGL(k_count_nnz<float>, grid, dim3(256), 0, st, (const float*)D, n, c.as<unsigned long long>());
  else
    hipLaunchKernelGGL(k_count_nnz<uint16_t>, grid, dim3(256), 0, st, (const uint16_t*)D, n,
                       c.as<unsigned long long>());
  FH_LAUNCH_CHECK();
  unsigned long long h = 0;
  FH_HIP(hipMemcpyAsync(&h, c.p, sizeof(h), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  return (int64_t)h;
}

}  // namespace fh
