// k_spmm.hip -- FP64 sparse x dense and dense x sparse products in the
// canonical order (one fma chain per output entry, k ascending).
//
// Layout: dense operands column-major with a leading dimension.  A wave holds
// 64 consecutive output rows of one output column, so the Y stores and (for
// dense x sparse) the A loads are coalesced; X gathers X[colind, j] stay in
// one column.  op(S) = S^T walks the CSC twin of S, whose rows are ascending
// inside each column, so the k order is the same as the oracle's.  CSR
// factors uploaded from the host with rows of at most 32 entries (and their
// transposes) are also kept as ELL views, read without the row pointer; the
// terms and their order are unchanged.
#include <algorithm>
#include <cstdlib>

#include "fh_internal.hpp"

namespace fh {

// Y(i,j) = alpha * sum_p v[p] * X(ci[p], j) + beta * Y(i,j)
// transX: X(k,j) is read as X[j + k*ldx] (op(X) = X^T).
// EW > 0: ci / v are the matrix's ELL view with EW slots per row.
template <bool TRANSX, int EW = 0>
__global__ __launch_bounds__(256) void k_csr_spmm_f64(int nrows, int ncols, const int* __restrict__ rp,
                                                       const int* __restrict__ ci, const double* __restrict__ v,
                                                       const double* __restrict__ X, int ldx,
                                                       double* __restrict__ Y, int ldy, double alpha, double beta) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= nrows || j >= ncols) return;
  double acc = 0.0;
  if (EW > 0) {
    // ELL view (EW slots per row, column -1 in unused slots; rp unused): the
    // same batches of 8 terms in the same order, without the row pointer
    const int* ec = ci + (size_t)i * EW;
    const double* ev = v + (size_t)i * EW;
    const double* xj = X + (size_t)j * ldx;
#pragma unroll
    for (int base = 0; base < EW; base += 8) {
      int cc[8];
      double ww[8], xx[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        cc[q] = ec[base + q];
        ww[q] = ev[base + q];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int c = max(cc[q], 0);
        xx[q] = TRANSX ? X[(size_t)c * ldx + j] : xj[c];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const double f = fma(ww[q], xx[q], acc);
        acc = cc[q] >= 0 ? f : acc;
      }
    }
    const double t = alpha * acc;
    double* y = Y + (size_t)j * ldy + i;
    *y = (beta == 0.0) ? t : fma(beta, *y, t);
    return;
  }
  const int b = rp[i], e = rp[i + 1];
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
  const int ew = (S.dtype == FH_F64) ? (tr ? S.csr_tell_w : S.csr_ell_w) : 0;
  const char* e_ell = getenv("FH_CSR_ELL");
  if (ew > 0 && ew % 8 == 0 && ew <= 32 && !(e_ell && e_ell[0] == '0')) {
    // ELL view: every row's slots at i * ew, no row pointer in the chain
    const int* eci = tr ? S.csr_tell_ci.as<int>() : S.csr_ell_ci.as<int>();
    const double* ev = tr ? S.csr_tell_v.as<double>() : S.csr_ell_v.as<double>();
#define FH_ELL64(T, W) \
  hipLaunchKernelGGL((k_csr_spmm_f64<T, W>), grid, dim3(256), 0, st, rows, ncols, nullptr, eci, ev, X, ldx, Y, ldy, alpha, beta)
    switch (ew * 2 + (transX ? 1 : 0)) {
      case 16: FH_ELL64(false, 8); break;
      case 17: FH_ELL64(true, 8); break;
      case 32: FH_ELL64(false, 16); break;
      case 33: FH_ELL64(true, 16); break;
      case 48: FH_ELL64(false, 24); break;
      case 49: FH_ELL64(true, 24); break;
      case 64: FH_ELL64(false, 32); break;
      default: FH_ELL64(true, 32); break;
    }
#undef FH_ELL64
    FH_LAUNCH_CHECK();
    return;
  }
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

// ------------------------------------------------------------ fused chain
// Small FP64 CSR chains in one launch: a workgroup owns WC columns of X and
// keeps its whole intermediate (rows x WC, FP64, column-major inside the
// workgroup) in LDS, double-buffered.  For every factor in turn it stages the
// factor's row pointers, column indices and values into LDS with coalesced
// loads, then each thread runs the fma chain of its rows from LDS; a barrier
// separates the phases.  Only the last factor writes Y.  Workgroups share
// nothing.  Each output entry is the fma chain of k_csr_spmm_f64 (k ascending
// from 0, then t = alpha*acc and beta == 0 ? t : fma(beta, y, t) on the last
// factor; intermediate factors have alpha = 1, beta = 0), so the result has
// the same bits as the per-factor launches.
constexpr int CHAIN_MAX = 16;
constexpr int CHAIN_LDS_BYTES = 160 * 1024;
struct ChainArgsF64 {
  ChainCsrOp f[CHAIN_MAX];
  int J, in_rows, R, max_nnz;
};


// Staged entry e sits at slot e + e/8: rows of ~8 entries would otherwise
// put the 64 lanes of a wave (one row each) on a handful of LDS banks.
__host__ __device__ __forceinline__ int chain_slot(int e) { return e + (e >> 3); }

// LDS carve-up (bytes): two activation buffers of R*WC doubles, then the
// staged factor: values (chain_slot(max_nnz) doubles), row pointers (R + 1
// ints), column indices (chain_slot(max_nnz) ints)
static size_t chain_lds_bytes(int R, int max_nnz, int wc) {
  const size_t slots = (size_t)chain_slot(max_nnz);
  return sizeof(double) * ((size_t)2 * R * wc + slots) + sizeof(int) * ((size_t)R + 1 + slots);
}

// Staged-factor share of one thread held in registers while the previous
// factor computes (prefetch): up to PF_NNZ entries and PF_RP row pointers.
constexpr int PF_NNZ = 16, PF_RP = 4;

constexpr int CHAIN_THREADS = 512;

template <int WC>
__global__ __launch_bounds__(CHAIN_THREADS) void k_chain_csr_f64(ChainArgsF64 a, const double* __restrict__ X, int ldx, int p,
                                                        double* __restrict__ Y, int ldy, double alpha, double beta) {
  extern __shared__ double lds[];
  const int t = threadIdx.x, nt = blockDim.x;
  const int c0 = blockIdx.x * WC;
  const int R = a.R;
  double* cur = lds;
  double* nxt = lds + (size_t)R * WC;
  double* sv = lds + (size_t)2 * R * WC;
  int* srp = reinterpret_cast<int*>(sv + chain_slot(a.max_nnz));
  int* sci = srp + R + 1;
  // the next factor's share of this thread, in registers (used when every
  // factor fits PF_NNZ / PF_RP entries per thread)
  const bool pf = a.max_nnz <= PF_NNZ * nt && R + 1 <= PF_RP * nt;
  int prp[PF_RP], pci[PF_NNZ];
  double pv[PF_NNZ];
  auto fetch = [&](const ChainCsrOp& S) {
#pragma unroll
    for (int u = 0; u < PF_RP; ++u) {
      const int i = t + u * nt;
      prp[u] = i <= S.rows ? S.rp[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < PF_NNZ; ++u) {
      const int i = t + u * nt;
      const bool on = i < S.nnz;
      pci[u] = on ? S.ci[i] : 0;
      pv[u] = on ? S.v[i] : 0.0;
    }
  };
  auto put = [&](const ChainCsrOp& S) {
#pragma unroll
    for (int u = 0; u < PF_RP; ++u) {
      const int i = t + u * nt;
      if (i <= S.rows) srp[i] = prp[u];
    }
#pragma unroll
    for (int u = 0; u < PF_NNZ; ++u) {
      const int i = t + u * nt;
      if (i < S.nnz) {
        sci[chain_slot(i)] = pci[u];
        sv[chain_slot(i)] = pv[u];
      }
    }
  };
  const int n0 = a.in_rows;
  if (pf) fetch(a.f[0]);
  for (int i = t; i < n0 * WC; i += nt) {
    const int c = i / n0, r = i - c * n0;
    cur[(size_t)c * R + r] = (c0 + c < p) ? X[(size_t)(c0 + c) * ldx + r] : 0.0;
  }
  for (int f = 0; f < a.J; ++f) {
    const ChainCsrOp S = a.f[f];
    if (pf) {
      put(S);
    } else {
      for (int i = t; i <= S.rows; i += nt) srp[i] = S.rp[i];
      for (int i = t; i < S.nnz; i += nt) {
        sci[chain_slot(i)] = S.ci[i];
        sv[chain_slot(i)] = S.v[i];
      }
    }
    __syncthreads();  // staged factor (and, for f > 0, the previous output) visible
    if (pf && f + 1 < a.J) fetch(a.f[f + 1]);  // in flight during this factor's rows
    const bool last = f == a.J - 1;
    for (int r = t; r < S.rows; r += nt) {
      const int b = srp[r], e = srp[r + 1];
      double acc[WC];
#pragma unroll
      for (int c = 0; c < WC; ++c) acc[c] = 0.0;
      // batches of 8 terms: index/value reads, then the gathers, then the fma
      // chain in k order (terms past the row end are not accumulated)
      for (int base = b; base < e; base += 8) {
        const int nb = min(8, e - base);
        int kk[8];
        double ww[8], xx[8][WC];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int pq = chain_slot(base + min(q, nb - 1));
          kk[q] = sci[pq];
          ww[q] = sv[pq];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int c = 0; c < WC; ++c) xx[q][c] = cur[(size_t)c * R + kk[q]];
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int c = 0; c < WC; ++c) {
            const double fq = fma(ww[q], xx[q][c], acc[c]);
            acc[c] = q < nb ? fq : acc[c];
          }
      }
      if (!last) {
#pragma unroll
        for (int c = 0; c < WC; ++c) nxt[(size_t)c * R + r] = acc[c];
      } else {
#pragma unroll
        for (int c = 0; c < WC; ++c) {
          if (c0 + c >= p) break;
          double* y = Y + (size_t)(c0 + c) * ldy + r;
          const double tt = alpha * acc[c];
          *y = (beta == 0.0) ? tt : fma(beta, *y, tt);
        }
      }
    }
    __syncthreads();  // every row done before the staging area and cur are reused
    double* sw = cur;
    cur = nxt;
    nxt = sw;
  }
}

// LDS bytes a fused-chain workgroup may use: the device's per-workgroup
// limit (160 KiB on gfx950), 64 KiB if the kernels cannot be given more.
// Queried once per process (one device).
static size_t chain_lds_budget() {
  static size_t budget = 0;
  if (budget) return budget;
  int dev = 0, v = 0;
  size_t b = 64 * 1024;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > (int)b) {
    const int want = std::min(v, CHAIN_LDS_BYTES);
    const bool ok =
        hipFuncSetAttribute((const void*)k_chain_csr_f64<1>, hipFuncAttributeMaxDynamicSharedMemorySize, want) ==
            hipSuccess &&
        hipFuncSetAttribute((const void*)k_chain_csr_f64<2>, hipFuncAttributeMaxDynamicSharedMemorySize, want) ==
            hipSuccess &&
        hipFuncSetAttribute((const void*)k_chain_csr_f64<4>, hipFuncAttributeMaxDynamicSharedMemorySize, want) ==
            hipSuccess &&
        hipFuncSetAttribute((const void*)k_chain_csr_f64<8>, hipFuncAttributeMaxDynamicSharedMemorySize, want) ==
            hipSuccess;
    if (ok) b = (size_t)want;
  }
  (void)hipGetLastError();  // a refused attribute leaves the 64 KiB budget, not an error
  budget = b;
  return budget;
}

// Columns per workgroup: FH_CHAIN_WC (1/2/4/8), else the narrowest (most
// workgroups; measured fastest) that keeps the grid within 1024 workgroups
// and whose LDS carve-up fits.
static int chain_wc(int R, int max_nnz, int p) {
  const size_t budget = chain_lds_budget();
  auto fits = [&](int w) { return chain_lds_bytes(R, max_nnz, w) <= budget; };
  if (const char* e = getenv("FH_CHAIN_WC")) {
    const int w = atoi(e);
    if ((w == 1 || w == 2 || w == 4 || w == 8) && fits(w)) return w;
  }
  int w = 1;
  while (w < 8 && (p + w - 1) / w > 1024) w *= 2;
  return fits(w) ? w : 0;
}

template <int WC>
static void chain_go(const ChainArgsF64& a, const double* X, int ldx, int p, double* Y, int ldy, double alpha,
                     double beta, hipStream_t st) {
  const size_t lds = chain_lds_bytes(a.R, a.max_nnz, WC);
  const dim3 grid((unsigned)((p + WC - 1) / WC));
  hipLaunchKernelGGL(k_chain_csr_f64<WC>, grid, dim3(CHAIN_THREADS), lds, st, a, X, ldx, p, Y, ldy, alpha, beta);
}

// Fused by default for chains of at least 6 factors.  Measured on 1024^2
// factors with 8 nnz per row, X 1024x64, per product: 2 factors 9.1 us fused
// vs 7.6 us factor by factor, 5: 19.0 vs 19.7, 10: 34.9 vs 39.8.  Inside the
// fused launch each factor still waits ~3 us for its staged row pointers,
// indices and values (a global-memory latency the one-factor-ahead prefetch
// only partly hides); a separate launch costs ~4 us.  FH_CHAIN_FUSED=1 / 0
// forces either way.
int csr_chain_plan_f64(const std::vector<ChainCsrOp>& ops, int in_rows, int p) {
  const int J = (int)ops.size();
  if (mul_mode() == FH_MUL_PER_FACTOR) return 0;
  const char* e = getenv("FH_CHAIN_FUSED");
  if (mul_mode() != FH_MUL_FUSED && e && e[0] == '0') return 0;
  const bool forced = mul_mode() == FH_MUL_FUSED || (e && e[0] == '1');
  if (J < 2 || J > CHAIN_MAX || p <= 0 || (!forced && J < 6)) return 0;
  // launches over ELL views (no row pointer in front of the gathers) beat the
  // one-launch chain at every length measured (5 to 10 factors of 1024^2)
  bool all_ell = true;
  for (const auto& o : ops) all_ell = all_ell && o.ell;
  if (!forced && all_ell) return 0;
  int R = in_rows, max_nnz = 0;
  for (const auto& o : ops) {
    R = std::max(R, o.rows);
    max_nnz = std::max(max_nnz, o.nnz);
  }
  if (R <= 0) return 0;
  const int wc = chain_wc(R, max_nnz, p);
  // one launch pays off while the chain is small next to the per-launch cost
  // of the factor-by-factor path: every workgroup walks every factor
  if (wc == 0 || (p + wc - 1) / wc > 1024) return 0;
  return wc;
}

void csr_chain_run_f64(const std::vector<ChainCsrOp>& ops, int in_rows, int wc, const double* X, int ldx, int p,
                       double* Y, int ldy, double alpha, double beta, hipStream_t st) {
  const int J = (int)ops.size();
  FH_CHECK(J >= 2 && J <= CHAIN_MAX && (wc == 1 || wc == 2 || wc == 4 || wc == 8), "fused chain: bad plan");
  ChainArgsF64 a{};
  a.J = J;
  a.in_rows = in_rows;
  a.R = in_rows;
  a.max_nnz = 0;
  for (int i = 0; i < J; ++i) {
    a.f[i] = ops[i];
    a.R = std::max(a.R, ops[i].rows);
    a.max_nnz = std::max(a.max_nnz, ops[i].nnz);
  }
  FH_CHECK(chain_lds_bytes(a.R, a.max_nnz, wc) <= chain_lds_budget(), "fused chain: LDS overflow");
  switch (wc) {
    case 8: chain_go<8>(a, X, ldx, p, Y, ldy, alpha, beta, st); break;
    case 4: chain_go<4>(a, X, ldx, p, Y, ldy, alpha, beta, st); break;
    case 2: chain_go<2>(a, X, ldx, p, Y, ldy, alpha, beta, st); break;
    default: chain_go<1>(a, X, ldx, p, Y, ldy, alpha, beta, st); break;
  }
  FH_LAUNCH_CHECK();
}

// ------------------------------------------------------ prefetched chain
// FP64 CSR chains whose factors and input have at most 1024 rows, in one
// launch of 1024 threads per workgroup: thread t owns row t of every factor,
// a workgroup WC columns of X.  Every factor's row pointers are copied into
// LDS at the start; a thread loads the first PF entries (values, columns) of
// its row of factor f + 2 while it computes factor f (two register sets,
// loads through buffer resources that return zeros past the factor's
// entries), and the barriers between factors order only the LDS activations,
// so those loads stay in flight across them.  Rows longer than PF read the
// rest of their entries in batches of 8.  Each output entry is the fma chain of
// k_csr_spmm_f64 (k ascending from 0; the last factor applies alpha / beta),
// so the bits are those of the factor-by-factor path.
constexpr int PFC_THREADS = 1024;
constexpr int PFC_RP_INTS = 16 * 1024;  // LDS ints for the row pointers of the whole chain

typedef unsigned int pfc_u4 __attribute__((ext_vector_type(4)));
typedef double pfc_d2 __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(1))) double* pfc_gd;
typedef const __attribute__((address_space(1))) int* pfc_gi;

template <int PF>
struct PfcRow {
  double v[PF];
  int c[PF];
  int p, n;  // first entry, count
};

__device__ __forceinline__ void pfc_lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// loads the first PF entries of row r of factor S (rows past the factor get
// n = 0 and read entries that are not used)
template <int PF>
__device__ __forceinline__ void pfc_fetch(const ChainCsrOp& S, const int* rpl, int rpo, int r, PfcRow<PF>& f) {
  const int rr = min(r, S.rows - 1);
  const int p0 = rpl[rpo + rr], p1 = rpl[rpo + rr + 1];
  f.p = p0;
  f.n = r < S.rows ? p1 - p0 : 0;
  const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc((void*)S.v, 0, S.nnz * 8, 0x00020000);
  const __amdgpu_buffer_rsrc_t rc = __builtin_amdgcn_make_buffer_rsrc((void*)S.ci, 0, S.nnz * 4, 0x00020000);
#pragma unroll
  for (int k = 0; k < PF; k += 2) {
    const pfc_d2 w = __builtin_bit_cast(pfc_d2, __builtin_amdgcn_raw_buffer_load_b128(rv, (p0 + k) * 8, 0, 0));
    f.v[k] = w.x;
    f.v[k + 1] = w.y;
  }
#pragma unroll
  for (int k = 0; k < PF; k += 4) {
    const pfc_u4 c = __builtin_amdgcn_raw_buffer_load_b128(rc, (p0 + k) * 4, 0, 0);
    f.c[k] = (int)c.x;
    f.c[k + 1] = (int)c.y;
    f.c[k + 2] = (int)c.z;
    f.c[k + 3] = (int)c.w;
  }
}

// the prefetched registers are read here on every path (see k_power.hip)
template <int PF>
__device__ __forceinline__ void pfc_settle(const PfcRow<PF>& f) {
#pragma unroll
  for (int k = 0; k < PF; ++k) asm volatile("" ::"v"(f.v[k]), "v"(f.c[k]));
}

// one factor: row t of S from cur (R x WC in LDS) into acc, entries from f
template <int PF, int WC>
__device__ __forceinline__ void pfc_row(const ChainCsrOp& S, const PfcRow<PF>& f, const double* cur, int R,
                                        double (&acc)[WC]) {
#pragma unroll
  for (int c = 0; c < WC; ++c) acc[c] = 0.0;
#pragma unroll
  for (int k = 0; k < PF; ++k)
    if (k < f.n) {
#pragma unroll
      for (int c = 0; c < WC; ++c) acc[c] = fma(f.v[k], cur[c * R + f.c[k]], acc[c]);
    }
  if (f.n > PF) {
    // the rest of the row in batches of 8: loads issued together (indices
    // clamped inside the row), then the fma chain in k order
    const pfc_gd v = (pfc_gd)S.v;
    const pfc_gi ci = (pfc_gi)S.ci;
    const int qe = f.p + f.n;
    for (int base = f.p + PF; base < qe; base += 8) {
      double w[8];
      int k[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int q = min(base + u, qe - 1);
        w[u] = v[q];
        k[u] = ci[q];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (base + u < qe) {
#pragma unroll
          for (int c = 0; c < WC; ++c) acc[c] = fma(w[u], cur[c * R + k[u]], acc[c]);
        }
    }
  }
}

template <int PF, int WC>
__device__ __forceinline__ void pfc_factor(const ChainArgsF64& a, const int* rpl, const int* rpo, int f,
                                           const double* cur, double* nxt, PfcRow<PF>& fr, int c0, int p,
                                           double* __restrict__ Y, int ldy, double alpha, double beta) {
  const int t = threadIdx.x;
  pfc_settle(fr);
  f = __builtin_amdgcn_readfirstlane(f);
  const ChainCsrOp S = a.f[f];
  double acc[WC];
  if (t < S.rows) pfc_row<PF, WC>(S, fr, cur, a.R, acc);
  // entries of the factor after next into the registers just consumed
  const int f2 = min(f + 2, a.J - 1);
  pfc_fetch<PF>(a.f[f2], rpl, rpo[f2], t, fr);
  if (t < S.rows) {
    if (f < a.J - 1) {
#pragma unroll
      for (int c = 0; c < WC; ++c) nxt[c * a.R + t] = acc[c];
    } else {
#pragma unroll
      for (int c = 0; c < WC; ++c)
        if (c0 + c < p) {
          double* y = Y + (size_t)(c0 + c) * ldy + t;
          const double tt = alpha * acc[c];
          *y = (beta == 0.0) ? tt : fma(beta, *y, tt);
        }
    }
  }
  pfc_lds_barrier();
}

template <int PF, int WC>
__global__ __launch_bounds__(PFC_THREADS) void k_chain_pf_f64(ChainArgsF64 a, const double* __restrict__ X, int ldx,
                                                              int p, double* __restrict__ Y, int ldy, double alpha,
                                                              double beta) {
  extern __shared__ double lds[];  // cur, nxt (R x WC each), then the row pointers
  __shared__ int rpo[CHAIN_MAX];
  const int t = threadIdx.x;
  const int c0 = blockIdx.x * WC;
  const int R = a.R;
  double* cur = lds;
  double* nxt = lds + (size_t)R * WC;
  int* rpl = reinterpret_cast<int*>(lds + (size_t)2 * R * WC);
  int off = 0;
  for (int f = 0; f < a.J; ++f) {
    const ChainCsrOp S = a.f[f];
    if (t == 0) rpo[f] = off;
    for (int i = t; i <= S.rows; i += PFC_THREADS) rpl[off + i] = S.rp[i];
    off += S.rows + 1;
  }
  const int n0 = a.in_rows;
  for (int i = t; i < n0 * WC; i += PFC_THREADS) {
    const int c = i / n0, r = i - c * n0;
    cur[(size_t)c * R + r] = (c0 + c < p) ? X[(size_t)(c0 + c) * ldx + r] : 0.0;
  }
  __syncthreads();
  PfcRow<PF> fa, fb;  // entries of the factors applied 1st, 3rd, ... and 2nd, 4th, ...
  pfc_fetch<PF>(a.f[0], rpl, rpo[0], t, fa);
  pfc_fetch<PF>(a.f[min(1, a.J - 1)], rpl, rpo[min(1, a.J - 1)], t, fb);
  for (int f = 0; f < a.J; f += 2) {
    pfc_factor<PF, WC>(a, rpl, rpo, f, cur, nxt, fa, c0, p, Y, ldy, alpha, beta);
    double* sw = cur;
    cur = nxt;
    nxt = sw;
    if (f + 1 == a.J) break;
    pfc_factor<PF, WC>(a, rpl, rpo, f + 1, cur, nxt, fb, c0, p, Y, ldy, alpha, beta);
    sw = cur;
    cur = nxt;
    nxt = sw;
  }
}

// Columns per workgroup of the prefetched chain: FH_CHAIN_PF_WC (1/2/4), else
// 1 up to 128 columns, 2 up to 512, 4 beyond (a workgroup reads every factor
// once whatever its width, so narrow panels spread that over more CUs while
// there are few columns).
static int pfc_wc(int p) {
  if (const char* e = getenv("FH_CHAIN_PF_WC")) {
    const int w = atoi(e);
    if (w == 1 || w == 2 || w == 4) return w;
  }
  return p <= 128 ? 1 : p <= 512 ? 2 : 4;
}

bool csr_chain_pf_fits(const std::vector<ChainCsrOp>& ops, int in_rows, int p) {
  const int J = (int)ops.size();
  if (J < 2 || J > CHAIN_MAX || p <= 0 || in_rows < 1 || in_rows > PFC_THREADS) return false;
  size_t rp_ints = 0;
  for (const auto& o : ops) {
    if (o.rows < 1 || o.rows > PFC_THREADS || (int64_t)o.nnz * 8 > INT32_MAX) return false;
    rp_ints += (size_t)o.rows + 1;
  }
  return rp_ints <= (size_t)PFC_RP_INTS;
}

int csr_chain_pf_default_wc(int p) { return pfc_wc(p); }

int csr_chain_pf_plan_f64(const std::vector<ChainCsrOp>& ops, int in_rows, int p) {
  if (mul_mode() != FH_MUL_AUTO) return 0;
  // FH_CHAIN_PF=0, or FH_CHAIN_FUSED set (either way: the staged one-launch
  // chain or the factor-by-factor launches are asked for), turns it off
  if (const char* e = getenv("FH_CHAIN_PF"))
    if (e[0] == '0') return 0;
  if (getenv("FH_CHAIN_FUSED")) return 0;
  return csr_chain_pf_fits(ops, in_rows, p) ? pfc_wc(p) : 0;
}

int csr_chain_fit_wc(const std::vector<ChainCsrOp>& ops, int in_rows, int p) {
  const int J = (int)ops.size();
  if (J < 2 || J > CHAIN_MAX || p <= 0) return 0;
  int R = in_rows, max_nnz = 0;
  for (const auto& o : ops) {
    R = std::max(R, o.rows);
    max_nnz = std::max(max_nnz, o.nnz);
  }
  if (R <= 0) return 0;
  const int wc = chain_wc(R, max_nnz, p);
  return (wc == 0 || (p + wc - 1) / wc > 1024) ? 0 : wc;
}

template <int PF, int WC>
static void pfc_go(const ChainArgsF64& a, size_t lds, const double* X, int ldx, int p, double* Y, int ldy,
                   double alpha, double beta, hipStream_t st) {
  static bool attr = false;
  if (!attr) {
    FH_HIP(hipFuncSetAttribute((const void*)k_chain_pf_f64<PF, WC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)(sizeof(double) * 2 * PFC_THREADS * 4 + sizeof(int) * PFC_RP_INTS)));
    attr = true;
  }
  const dim3 grid((unsigned)((p + WC - 1) / WC));
  hipLaunchKernelGGL((k_chain_pf_f64<PF, WC>), grid, dim3(PFC_THREADS), lds, st, a, X, ldx, p, Y, ldy, alpha, beta);
}

void csr_chain_pf_run_f64(const std::vector<ChainCsrOp>& ops, int in_rows, int wc, int pf, const double* X, int ldx,
                          int p, double* Y, int ldy, double alpha, double beta, hipStream_t st) {
  const int J = (int)ops.size();
  FH_CHECK(J >= 2 && J <= CHAIN_MAX && (wc == 1 || wc == 2 || wc == 4) && (pf == 8 || pf == 16),
           "prefetched chain: bad plan");
  ChainArgsF64 a{};
  a.J = J;
  a.in_rows = in_rows;
  a.R = in_rows;
  size_t rp_ints = 0;
  for (int i = 0; i < J; ++i) {
    a.f[i] = ops[i];
    a.R = std::max(a.R, ops[i].rows);
    rp_ints += (size_t)ops[i].rows + 1;
  }
  FH_CHECK(a.R <= PFC_THREADS && rp_ints <= (size_t)PFC_RP_INTS, "prefetched chain: shape out of range");
  const size_t lds = sizeof(double) * 2 * (size_t)a.R * wc + sizeof(int) * rp_ints;
#define FH_PFC(PF, WC) pfc_go<PF, WC>(a, lds, X, ldx, p, Y, ldy, alpha, beta, st)
  if (pf == 8) {
    if (wc == 1) FH_PFC(8, 1); else if (wc == 2) FH_PFC(8, 2); else FH_PFC(8, 4);
  } else {
    if (wc == 1) FH_PFC(16, 1); else if (wc == 2) FH_PFC(16, 2); else FH_PFC(16, 4);
  }
#undef FH_PFC
  FH_LAUNCH_CHECK();
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
