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
__global__ __launch_bounds__(PW_THREADS) void k_power_iteration(const PwOp* __restrict__ ops, int nops, int n,
                                                                 const double* __restrict__ x0, int maxit,
                                                                 double thr, double* __restrict__ out) {
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

double power_iteration_f64(const std::vector<ChainOp>& ops, int n, int maxit, double thr, int* flag,
                           hipStream_t st) {
  FH_CHECK(maxit > 0, "power_iteration: nbr_iter_max <= 0");
  FH_CHECK(!ops.empty(), "power_iteration: empty chain");
  std::vector<PwOp> h(ops.size());
  for (size_t s = 0; s < ops.size(); ++s) {
    const DMat& M = *ops[s].M;
    const bool tr = ops[s].op != 'N';
    PwOp& o = h[s];
    o.rows = tr ? M.n : M.m;
    o.cols = tr ? M.m : M.n;
    FH_CHECK(o.rows <= PW_VMAX && o.cols <= PW_VMAX, "power_iteration: dimension above 2048");
    if (M.kind == K_CSR) {
      o.kind = 1;
      o.trans = 0;
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
    }
  }
  const std::vector<double>& x0 = power_x0(n);
  DBuf dops(sizeof(PwOp) * h.size()), dx(sizeof(double) * n), dout(sizeof(double) * 3);
  FH_HIP(hipMemcpyAsync(dops.p, h.data(), sizeof(PwOp) * h.size(), hipMemcpyHostToDevice, st));
  FH_HIP(hipMemcpyAsync(dx.p, x0.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_power_iteration, dim3(1), dim3(PW_THREADS), 0, st, dops.as<PwOp>(), (int)h.size(), n,
                     dx.as<double>(), maxit, thr, dout.as<double>());
  FH_LAUNCH_CHECK();
  double res[3];
  FH_HIP(hipMemcpyAsync(res, dout.p, sizeof(res), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  if (res[2] == 1.0) throw Error(FH_ERR_NUMERIC, "Can't normalize a zero-norm vector.");
  if (flag) *flag = (int)res[1];
  return res[0];
}

}  // namespace fh
