// k_power.hip -- power iteration on a factor chain in one persistent launch.
//
// The reference (power_iteration, faust_linear_algebra.hpp:808-898) runs up
// to 100 dependent rounds of "normalise, apply the F F^H chain, Rayleigh
// quotient", each a string of tiny mat-vecs.  Here a single 1024-thread
// workgroup keeps both vectors in LDS and runs every round, every mat-vec and
// the stopping test on the device: no launch or host round trip per step.
// Mat-vecs are one fma chain per output row (k ascending); norms and dots use
// the canonical 256-lane chunk reduction, so the result matches the oracle
// bit for bit.
#include <cstdlib>
#include <map>
#include <mutex>

#include "fh_internal.hpp"

namespace fh {

constexpr int PW_THREADS = 1024;
constexpr int PW_VMAX = 2048;  // max vector length kept in LDS (3 vectors: 48 KiB)

struct PwOp {
  int kind;   // 0 dense, 1 row-compressed sparse
  int trans;  // dense only: op(A) = A^T
  int rows, cols;
  const int* rp;
  const int* ci;
  const double* v;  // sparse values or dense data (ld = rows of A)
  int lda;
  int64_t nnz;  // sparse only
};

// canonical reduction of dot(x, y) over n values held in LDS; every thread
// returns the same value.  red: 256 doubles of LDS scratch; tot: 1 double.
__device__ double pw_dot(const double* x, const double* y, int n, double* red, double* tot) {
  const int t = threadIdx.x;
  const int nch = (n + RED_CHUNK - 1) / RED_CHUNK;
  if (t == 0) *tot = 0.0;
  __syncthreads();
  for (int c = 0; c < nch; ++c) {
    if (t < RED_LANES) {
      double acc = 0.0;
      for (int m = 0; m < RED_CHUNK / RED_LANES; ++m) {
        int i = c * RED_CHUNK + m * RED_LANES + t;
        if (i < n) acc = fma(x[i], y[i], acc);
      }
      red[t] = acc;
    }
    __syncthreads();
    for (int s = RED_LANES / 2; s >= 1; s >>= 1) {
      if (t < s) red[t] = red[t] + red[t + s];
      __syncthreads();
    }
    if (t == 0) *tot = *tot + red[0];
    __syncthreads();
  }
  double r = *tot;
  __syncthreads();
  return r;
}

// out[0] = lambda, out[1] = flag (iterations, or -1), out[2] = status
// (0 ok, 1 zero-norm vector, 2 A x == 0)
// One chain per workgroup: workgroup b runs jobs[b] (independent chains, e.g.
// the L and R norms of one PALM factor update, share a launch).
struct PwJob {
  const PwOp* ops;
  int nops;
  int n;
  const double* x0;
  double* out;
};

__global__ __launch_bounds__(PW_THREADS) void k_power_iteration(const PwJob* __restrict__ jobs, int maxit,
                                                                 double thr) {
  const PwJob job = jobs[blockIdx.x];
  const PwOp* __restrict__ ops = job.ops;
  const int nops = job.nops, n = job.n;
  const double* __restrict__ x0 = job.x0;
  double* __restrict__ out = job.out;
  __shared__ double xa[PW_VMAX], xb[PW_VMAX], xn[PW_VMAX];
  __shared__ double red[RED_LANES];
  __shared__ double tot;
  const int t = threadIdx.x;
  for (int i = t; i < n; i += PW_THREADS) xa[i] = x0[i];
  __syncthreads();
  double lambda_old = 1.0, lambda = 0.0;
  int it = 0;
  int status = 0;
  const double athr = fabs(thr);
  while ((fabs(lambda_old - lambda) > athr || fabs(lambda) <= athr) && it < maxit) {
    it++;
    lambda_old = lambda;
    double nrm = sqrt(pw_dot(xa, xa, n, red, &tot));
    if (nrm == 0.0) {
      status = 1;
      break;
    }
    const double inv = 1 / nrm;
    for (int i = t; i < n; i += PW_THREADS) xn[i] = xa[i] * inv;
    __syncthreads();
    // apply ops last to first: cur -> nxt
    const double* cur = xn;
    double* nxt = xb;
    int len = n;
    for (int s = nops - 1; s >= 0; --s) {
      const PwOp o = ops[s];
      for (int r = t; r < o.rows; r += PW_THREADS) {
        double acc = 0.0;
        if (o.kind == 1) {
          for (int p = o.rp[r]; p < o.rp[r + 1]; ++p) acc = fma(o.v[p], cur[o.ci[p]], acc);
        } else if (!o.trans) {
          const double* a = o.v + r;
          for (int kk = 0; kk < o.cols; ++kk) acc = fma(a[(size_t)kk * o.lda], cur[kk], acc);
        } else {
          const double* a = o.v + (size_t)r * o.lda;
          for (int kk = 0; kk < o.cols; ++kk) acc = fma(a[kk], cur[kk], acc);
        }
        nxt[r] = acc;
      }
      len = o.rows;
      __syncthreads();
      // ping-pong between xb and xa (xn stays the normalised start vector)
      cur = nxt;
      nxt = (nxt == xb) ? xa : xb;
    }
    // result vector is `cur`; move it into xa if needed
    if (cur != xa) {
      for (int i = t; i < len; i += PW_THREADS) xa[i] = cur[i];
      __syncthreads();
    }
    int nz = 0;
    for (int i = t; i < n; i += PW_THREADS) nz |= (xa[i] != 0.0);
    if (!__syncthreads_or(nz)) {
      status = 2;
      break;
    }
    lambda = pw_dot(xn, xa, n, red, &tot);
  }
  if (t == 0) {
    out[0] = (status == 2) ? 0.0 : lambda;
    out[1] = (it < maxit) ? (double)it : -1.0;
    out[2] = (double)status;
  }
}

// ---------------------------------------------------------------------------
// Wide form: one launch per operator per round, for chains whose operators are
// too large for a single CU to stream every round.  Each row is still one fma
// chain on one lane (k ascending), so the result is bit-identical to the
// persistent kernel; the row data is staged through LDS by all 256 threads of
// the workgroup in contiguous chunks, and the first R lanes (one per row) run
// the chains.  A device state carries lambda, the round count and a done flag: the
// host enqueues rounds ahead and launches after convergence return at once.
struct PwState {
  double lambda, lambda_old;
  int it, status, done, pad;
};

constexpr int PW_WIDE_THREADS = 256;
constexpr int PW_CH = 2048;  // entries per LDS chunk

// sparse rows (row pointers rp, indices ci, values v): rows r0..r0+R-1 per workgroup
__global__ __launch_bounds__(PW_WIDE_THREADS) void k_pw_csr(const int* __restrict__ rp, const int* __restrict__ ci,
                                                            const double* __restrict__ v, int rows, int cols, int R,
                                                            const double* __restrict__ cur, double* __restrict__ nxt,
                                                            const PwState* __restrict__ s) {
  __shared__ double xs[PW_VMAX];
  __shared__ double sv[PW_CH];
  __shared__ int sci[PW_CH];
  if (s->done) return;
  const int t = threadIdx.x;
  const int r0 = blockIdx.x * R, r1 = min(rows, r0 + R), nr = r1 - r0;
  for (int i = t; i < cols; i += PW_WIDE_THREADS) xs[i] = cur[i];
  const int p0 = rp[r0], p1 = rp[r1];
  int p = 0, pe = 0;
  if (t < nr) {
    p = rp[r0 + t];
    pe = rp[r0 + t + 1];
  }
  double acc = 0.0;
  for (int c0 = p0; c0 < p1; c0 += PW_CH) {
    const int c1 = min(p1, c0 + PW_CH);
    __syncthreads();  // previous chunk consumed; xs staged before the first
    for (int q = c0 + t; q < c1; q += PW_WIDE_THREADS) {
      sv[q - c0] = v[q];
      sci[q - c0] = ci[q];
    }
    __syncthreads();
    if (t < nr) {
      const int e = min(pe, c1);
      for (; p < e; ++p) acc = fma(sv[p - c0], xs[sci[p - c0]], acc);
    }
  }
  if (t < nr) nxt[r0 + t] = acc;
}

// dense op(A) rows (A column-major, leading dimension lda); trans: op(A) = A^T
__global__ __launch_bounds__(PW_WIDE_THREADS) void k_pw_dense(const double* __restrict__ A, int lda, int trans,
                                                              int rows, int cols, int R,
                                                              const double* __restrict__ cur,
                                                              double* __restrict__ nxt,
                                                              const PwState* __restrict__ s) {
  __shared__ double xs[PW_VMAX];
  __shared__ double tile[PW_CH];  // R rows x C columns, row stride C + 1 (odd: no bank conflicts)
  if (s->done) return;
  const int t = threadIdx.x;
  const int r0 = blockIdx.x * R, nr = min(rows, r0 + R) - r0;
  const int C = PW_CH / R - 2;
  for (int i = t; i < cols; i += PW_WIDE_THREADS) xs[i] = cur[i];
  double acc = 0.0;
  for (int c0 = 0; c0 < cols; c0 += C) {
    const int nc = min(C, cols - c0);
    __syncthreads();
    if (trans) {  // row i of A^T = column r0+i of A: contiguous in c
      for (int q = t; q < nr * nc; q += PW_WIDE_THREADS) {
        const int i = q / nc, c = q - i * nc;
        tile[i * (C + 1) + c] = A[(size_t)(r0 + i) * lda + c0 + c];
      }
    } else {  // row i of A: element (i, c) at (c0+c)*lda + r0 + i
      for (int q = t; q < nr * nc; q += PW_WIDE_THREADS) {
        const int c = q / nr, i = q - c * nr;
        tile[i * (C + 1) + c] = A[(size_t)(c0 + c) * lda + r0 + i];
      }
    }
    __syncthreads();
    if (t < nr) {
      const double* row = tile + t * (C + 1);
      for (int c = 0; c < nc; ++c) acc = fma(row[c], xs[c0 + c], acc);
    }
  }
  if (t < nr) nxt[r0 + t] = acc;
}

// Operators with more than PW_VMAX columns: x read from global memory (the
// L2 holds it), one row per thread, the same fma chain per row.
__global__ __launch_bounds__(PW_WIDE_THREADS) void k_pw_csr_g(const int* __restrict__ rp, const int* __restrict__ ci,
                                                              const double* __restrict__ v, int rows,
                                                              const double* __restrict__ cur,
                                                              double* __restrict__ nxt,
                                                              const PwState* __restrict__ s) {
  if (s->done) return;
  const int r = blockIdx.x * PW_WIDE_THREADS + threadIdx.x;
  if (r >= rows) return;
  double acc = 0.0;
  for (int p = rp[r], pe = rp[r + 1]; p < pe; ++p) acc = fma(v[p], cur[ci[p]], acc);
  nxt[r] = acc;
}

__global__ __launch_bounds__(PW_WIDE_THREADS) void k_pw_dense_g(const double* __restrict__ A, int lda, int trans,
                                                                int rows, int cols, const double* __restrict__ cur,
                                                                double* __restrict__ nxt,
                                                                const PwState* __restrict__ s) {
  if (s->done) return;
  const int r = blockIdx.x * PW_WIDE_THREADS + threadIdx.x;
  if (r >= rows) return;
  double acc = 0.0;
  if (trans) {  // row r of A^T = column r of A
    const double* col = A + (size_t)r * lda;
    for (int c = 0; c < cols; ++c) acc = fma(col[c], cur[c], acc);
  } else {
    for (int c = 0; c < cols; ++c) acc = fma(A[(size_t)c * lda + r], cur[c], acc);
  }
  nxt[r] = acc;
}

// One round boundary.  mode 1: finish the previous round (xa = chain output):
// zero test and lambda = <xn, xa>.  Then the loop test of the persistent
// kernel, and if it continues: it++, normalise xa into xn.
__global__ __launch_bounds__(PW_THREADS) void k_pw_step(int n, double* __restrict__ xa, double* __restrict__ xn,
                                                         PwState* __restrict__ s, int maxit, double thr, int mode) {
  __shared__ double red[RED_LANES];
  __shared__ double tot;
  const int t = threadIdx.x;
  const PwState st = *s;
  __syncthreads();
  if (st.done) return;
  double lambda = st.lambda, lambda_old = st.lambda_old;
  int it = st.it;
  if (mode == 1) {
    int nz = 0;
    for (int i = t; i < n; i += PW_THREADS) nz |= (xa[i] != 0.0);
    if (!__syncthreads_or(nz)) {
      if (t == 0) {
        s->status = 2;
        s->done = 1;
      }
      return;
    }
    lambda = pw_dot(xn, xa, n, red, &tot);
  }
  const double athr = fabs(thr);
  if (!((fabs(lambda_old - lambda) > athr || fabs(lambda) <= athr) && it < maxit)) {
    if (t == 0) {
      s->lambda = lambda;
      s->done = 1;
    }
    return;
  }
  it++;
  lambda_old = lambda;
  const double nrm = sqrt(pw_dot(xa, xa, n, red, &tot));
  if (nrm == 0.0) {
    if (t == 0) {
      s->it = it;
      s->status = 1;
      s->done = 1;
    }
    return;
  }
  const double inv = 1 / nrm;
  for (int i = t; i < n; i += PW_THREADS) xn[i] = xa[i] * inv;
  if (t == 0) {
    s->lambda = lambda;
    s->lambda_old = lambda_old;
    s->it = it;
  }
}

__global__ void k_pw_init(PwState* __restrict__ s) {
  if (threadIdx.x == 0) {
    s->lambda = 0.0;
    s->lambda_old = 1.0;
    s->it = 0;
    s->status = 0;
    s->done = 0;
    s->pad = 0;
  }
}

const std::vector<double>& power_x0(int n) {
  static std::mutex mu;
  static std::map<int, std::vector<double>> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(n);
  if (it != cache.end()) return it->second;
  // srand(0xF4+'u'+57); Eigen::Matrix::Random: x + (y-x)*Scalar(std::rand())/Scalar(RAND_MAX) on [-1, 1]
  srand(0xF4 + 'u' + 57);
  std::vector<double> x(n);
  for (int i = 0; i < n; ++i) x[i] = -1.0 + (2.0 * (double)std::rand()) / (double)RAND_MAX;
  return cache.emplace(n, std::move(x)).first->second;
}

// Per-round time estimates (seconds): the persistent kernel streams every
// operator through one CU (~100 GB/s); the wide form pays a launch per
// operator but spreads the rows over the chip, leaving each row's fma chain
// (about 4 ns per term) as its floor.  FH_POWER_WIDE=0/1 forces a form.
static bool use_wide_form(const std::vector<PwOp>& h) {
  if (const char* e = getenv("FH_POWER_WIDE")) {
    if (e[0] == '0') return false;
    if (e[0] == '1') return true;
  }
  double single = 3e-6, wide = 4e-6;
  for (const PwOp& o : h) {
    const double entries = o.kind == 1 ? (double)o.nnz : (double)o.rows * o.cols;
    const double bytes = entries * (o.kind == 1 ? 12.0 : 8.0);
    single += bytes / 100e9 + 0.4e-6;
    wide += 3e-6 + (entries / std::max(o.rows, 1)) * 4e-9 + bytes / 1.5e12;
  }
  return wide < single;
}

static double power_iteration_wide(const std::vector<PwOp>& h, int n, const std::vector<double>& x0, int maxit,
                                   double thr, int* flag, hipStream_t st) {
  int vmax = n;
  for (const PwOp& o : h) vmax = std::max(vmax, std::max(o.rows, o.cols));
  DBuf xa(sizeof(double) * n), xn(sizeof(double) * n), b0(sizeof(double) * vmax), b1(sizeof(double) * vmax),
      ds(sizeof(PwState));
  PwState* dst = ds.as<PwState>();
  FH_HIP(hipMemcpyAsync(xa.p, x0.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_pw_init, dim3(1), dim3(64), 0, st, dst);
  FH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_pw_step, dim3(1), dim3(PW_THREADS), 0, st, n, xa.as<double>(), xn.as<double>(), dst, maxit,
                     thr, 0);
  FH_LAUNCH_CHECK();
  const int nops = (int)h.size();
  PwState hs{};
  int next_check = 4;
  for (int round = 1; round <= maxit; ++round) {
    const double* cur = xn.as<double>();
    for (int s = nops - 1; s >= 0; --s) {
      const PwOp& o = h[s];
      double* out = (s == 0) ? xa.as<double>() : (((nops - 1 - s) & 1) ? b1.as<double>() : b0.as<double>());
      int R = 1;
      while (R < 8 && (o.rows + 2 * R - 1) / (2 * R) >= 256) R *= 2;
      const dim3 grid((o.rows + R - 1) / R);
      if (o.cols > PW_VMAX) {  // x does not fit LDS: read it from global memory
        const dim3 gg((o.rows + PW_WIDE_THREADS - 1) / PW_WIDE_THREADS);
        if (o.kind == 1)
          hipLaunchKernelGGL(k_pw_csr_g, gg, dim3(PW_WIDE_THREADS), 0, st, o.rp, o.ci, o.v, o.rows, cur, out,
                             (const PwState*)dst);
        else
          hipLaunchKernelGGL(k_pw_dense_g, gg, dim3(PW_WIDE_THREADS), 0, st, o.v, o.lda, o.trans, o.rows, o.cols,
                             cur, out, (const PwState*)dst);
      } else if (o.kind == 1)
        hipLaunchKernelGGL(k_pw_csr, grid, dim3(PW_WIDE_THREADS), 0, st, o.rp, o.ci, o.v, o.rows, o.cols, R, cur,
                           out, (const PwState*)dst);
      else
        hipLaunchKernelGGL(k_pw_dense, grid, dim3(PW_WIDE_THREADS), 0, st, o.v, o.lda, o.trans, o.rows, o.cols, R,
                           cur, out, (const PwState*)dst);
      FH_LAUNCH_CHECK();
      cur = out;
    }
    hipLaunchKernelGGL(k_pw_step, dim3(1), dim3(PW_THREADS), 0, st, n, xa.as<double>(), xn.as<double>(), dst,
                       maxit, thr, 1);
    FH_LAUNCH_CHECK();
    if (round == next_check || round == maxit) {
      FH_HIP(hipMemcpyAsync(&hs, dst, sizeof(PwState), hipMemcpyDeviceToHost, st));
      FH_HIP(hipStreamSynchronize(st));
      if (hs.done) break;
      next_check += 8;
    }
  }
  if (!hs.done) {  // every round ran: one more step evaluates the loop test
    hipLaunchKernelGGL(k_pw_step, dim3(1), dim3(PW_THREADS), 0, st, n, xa.as<double>(), xn.as<double>(), dst,
                       maxit, thr, 1);
    FH_LAUNCH_CHECK();
    FH_HIP(hipMemcpyAsync(&hs, dst, sizeof(PwState), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
  }
  if (hs.status == 1) throw Error(FH_ERR_NUMERIC, "Can't normalize a zero-norm vector.");
  if (flag) *flag = (hs.it < maxit) ? hs.it : -1;
  return hs.status == 2 ? 0.0 : hs.lambda;
}

static std::vector<PwOp> pw_ops(const std::vector<ChainOp>& ops) {
  FH_CHECK(!ops.empty(), "power_iteration: empty chain");
  std::vector<PwOp> h(ops.size());
  for (size_t s = 0; s < ops.size(); ++s) {
    const DMat& M = *ops[s].M;
    const bool tr = ops[s].op != 'N';
    PwOp& o = h[s];
    o.rows = tr ? M.n : M.m;
    o.cols = tr ? M.m : M.n;
    if (M.kind == K_CSR) {
      FH_CHECK(M.dtype == FH_F64, "power_iteration: FP64 dense or CSR factors only");
      o.kind = 1;
      o.trans = 0;
      o.nnz = M.nnz;
      if (tr) {
        FH_CHECK(M.has_csc, "power_iteration: CSC twin missing");
        o.rp = M.cp.as<int>();
        o.ci = M.ri.as<int>();
        o.v = M.cv.as<double>();
      } else {
        o.rp = M.rp.as<int>();
        o.ci = M.ci.as<int>();
        o.v = M.v.as<double>();
      }
    } else {
      FH_CHECK(M.kind == K_DENSE && M.dtype == FH_F64, "power_iteration: FP64 dense or CSR factors only");
      o.kind = 0;
      o.trans = tr ? 1 : 0;
      o.rp = o.ci = nullptr;
      o.v = M.d.as<double>();
      o.lda = M.m;
      o.nnz = 0;
    }
  }
  return h;
}

// the persistent kernel keeps every vector in LDS: up to PW_VMAX entries
static bool pw_persistent(const std::vector<PwOp>& h, int n) {
  bool fits = n <= PW_VMAX;
  for (const PwOp& o : h) fits = fits && o.rows <= PW_VMAX && o.cols <= PW_VMAX;
  return fits && !use_wide_form(h);
}

// chains[k] (ops, n) in one launch, one workgroup each; lam[k] and flags[k]
// receive each chain's result.  A zero-norm start vector throws, for the
// first such chain in order.
static void power_iteration_persistent(const std::vector<std::vector<PwOp>>& hs, const std::vector<int>& ns,
                                       int maxit, double thr, double* lam, int* flags, hipStream_t st) {
  const int J = (int)hs.size();
  size_t nops = 0, nx = 0;
  for (int k = 0; k < J; ++k) {
    nops += hs[k].size();
    nx += ns[k];
  }
  DBuf dops(sizeof(PwOp) * nops), dx(sizeof(double) * nx), dout(sizeof(double) * 3 * J), djobs(sizeof(PwJob) * J);
  std::vector<PwOp> allops;
  std::vector<PwJob> jobs(J);
  size_t oo = 0, ox = 0;
  for (int k = 0; k < J; ++k) {
    allops.insert(allops.end(), hs[k].begin(), hs[k].end());
    FH_HIP(hipMemcpyAsync(dx.as<double>() + ox, power_x0(ns[k]).data(), sizeof(double) * ns[k],
                          hipMemcpyHostToDevice, st));
    jobs[k] = PwJob{dops.as<PwOp>() + oo, (int)hs[k].size(), ns[k], dx.as<double>() + ox, dout.as<double>() + 3 * k};
    oo += hs[k].size();
    ox += ns[k];
  }
  FH_HIP(hipMemcpyAsync(dops.p, allops.data(), sizeof(PwOp) * nops, hipMemcpyHostToDevice, st));
  FH_HIP(hipMemcpyAsync(djobs.p, jobs.data(), sizeof(PwJob) * J, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_power_iteration, dim3(J), dim3(PW_THREADS), 0, st, djobs.as<PwJob>(), maxit, thr);
  FH_LAUNCH_CHECK();
  std::vector<double> res(3 * J);
  FH_HIP(hipMemcpyAsync(res.data(), dout.p, sizeof(double) * 3 * J, hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < J; ++k) {
    if (res[3 * k + 2] == 1.0) throw Error(FH_ERR_NUMERIC, "Can't normalize a zero-norm vector.");
    lam[k] = res[3 * k];
    flags[k] = (int)res[3 * k + 1];
  }
}

double power_iteration_f64(const std::vector<ChainOp>& ops, int n, int maxit, double thr, int* flag,
                           hipStream_t st) {
  FH_CHECK(maxit > 0, "power_iteration: nbr_iter_max <= 0");
  std::vector<PwOp> h = pw_ops(ops);
  if (!pw_persistent(h, n)) return power_iteration_wide(h, n, power_x0(n), maxit, thr, flag, st);
  double lam;
  int fl;
  power_iteration_persistent({h}, {n}, maxit, thr, &lam, &fl, st);
  if (flag) *flag = fl;
  return lam;
}

void power_iteration_pair_f64(const std::vector<ChainOp>& ops1, int n1, const std::vector<ChainOp>& ops2, int n2,
                              int maxit, double thr, double* lam1, double* lam2, int* flag, hipStream_t st) {
  FH_CHECK(maxit > 0, "power_iteration: nbr_iter_max <= 0");
  std::vector<PwOp> h1 = pw_ops(ops1), h2 = pw_ops(ops2);
  if (!pw_persistent(h1, n1) || !pw_persistent(h2, n2)) {
    *lam1 = power_iteration_f64(ops1, n1, maxit, thr, flag, st);
    *lam2 = power_iteration_f64(ops2, n2, maxit, thr, flag, st);
    return;
  }
  double lam[2];
  int fl[2];
  power_iteration_persistent({h1, h2}, {n1, n2}, maxit, thr, lam, fl, st);
  *lam1 = lam[0];
  *lam2 = lam[1];
  if (flag) *flag = fl[1];
}

}  // namespace fh
