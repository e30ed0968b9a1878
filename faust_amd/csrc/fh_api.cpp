// fh_api.cpp -- C ABI: matrix handles, products, Faust chains, projectors.
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <cstring>
#include <vector>

#include "fh_api_common.hpp"

using namespace fh;

namespace fh {

// Host-side CSR -> CSC (counting sort by column, stable in row order, so rows
// ascend inside each column).
static void host_csc(int m, int n, const int32_t* rp, const int32_t* ci, const void* vals, size_t esz,
                     std::vector<int>& cp, std::vector<int>& ri, std::vector<char>& cv) {
  const int64_t nnz = rp[m];
  cp.assign(n + 1, 0);
  ri.resize(nnz);
  cv.resize(nnz * esz);
  for (int64_t p = 0; p < nnz; ++p) cp[ci[p] + 1]++;
  for (int j = 0; j < n; ++j) cp[j + 1] += cp[j];
  std::vector<int> fill(cp.begin(), cp.end() - 1);
  const char* v = static_cast<const char*>(vals);
  for (int i = 0; i < m; ++i)
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
      int q = fill[ci[p]]++;
      ri[q] = i;
      std::memcpy(&cv[q * esz], v + p * esz, esz);
    }
}

DMatP make_csr_from_host(int dtype, int m, int n, int nnz, const int32_t* rp, const int32_t* ci, const void* vals,
                         hipStream_t st) {
  FH_CHECK(m >= 0 && n >= 0 && nnz >= 0, "csr: negative size");
  FH_CHECK(rp[0] == 0 && rp[m] == nnz, "csr: rowptr inconsistent with nnz");
  // columns must be sorted and in range inside each row (canonical k order)
  for (int i = 0; i < m; ++i) {
    FH_CHECK(rp[i] <= rp[i + 1], "csr: rowptr not non-decreasing");
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
      FH_CHECK(ci[p] >= 0 && ci[p] < n, "csr: column index out of range");
      FH_CHECK(p == rp[i] || ci[p] > ci[p - 1], "csr: column indices must be strictly increasing in each row");
    }
  }
  const size_t esz = dtype_size(dtype);
  auto M = std::make_shared<DMat>();
  M->kind = K_CSR;
  M->dtype = dtype;
  M->m = m;
  M->n = n;
  M->nnz = nnz;
  std::vector<int> cp, ri;
  std::vector<char> cv;
  host_csc(m, n, rp, ci, vals, esz, cp, ri, cv);
  M->rp.alloc(sizeof(int) * (m + 1));
  M->ci.alloc(sizeof(int) * std::max(nnz, 1));
  M->v.alloc(esz * std::max(nnz, 1));
  M->cp.alloc(sizeof(int) * (n + 1));
  M->ri.alloc(sizeof(int) * std::max(nnz, 1));
  M->cv.alloc(esz * std::max(nnz, 1));
  FH_HIP(hipMemcpyAsync(M->rp.p, rp, sizeof(int) * (m + 1), hipMemcpyHostToDevice, st));
  FH_HIP(hipMemcpyAsync(M->cp.p, cp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, st));
  if (nnz > 0) {
    FH_HIP(hipMemcpyAsync(M->ci.p, ci, sizeof(int) * nnz, hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(M->v.p, vals, esz * nnz, hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(M->ri.p, ri.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(M->cv.p, cv.data(), esz * nnz, hipMemcpyHostToDevice, st));
  }
  M->has_csc = true;
  M->max_row = 0;
  for (int i = 0; i < m; ++i) M->max_row = std::max(M->max_row, rp[i + 1] - rp[i]);
  M->max_col = 0;
  for (int j = 0; j < n; ++j) M->max_col = std::max(M->max_col, cp[j + 1] - cp[j]);
  if (dtype == FH_F32 || dtype == FH_F64) {
    // ELL views for the gather kernels (rows of S, rows of S^T): the widest
    // row rounded up to the kernel's batch, at most 16 slots for FP32 (in 4s)
    // and 32 for FP64 (in 8s)
    const int quant = dtype == FH_F32 ? 4 : 8, wmax = dtype == FH_F32 ? 16 : 32;
    auto ell = [&](int nr, const int* ptr, const int* idx, const char* val, int widest, int& w, DBuf& dc, DBuf& dv) {
      w = 0;
      if (nr == 0 || widest > wmax) return;
      w = std::max(quant, (widest + quant - 1) / quant * quant);
      std::vector<int> hc((size_t)nr * w, -1);
      std::vector<char> hv((size_t)nr * w * esz, 0);
      for (int r = 0; r < nr; ++r)
        for (int e = ptr[r]; e < ptr[r + 1]; ++e) {
          hc[(size_t)r * w + (e - ptr[r])] = idx[e];
          std::memcpy(hv.data() + ((size_t)r * w + (e - ptr[r])) * esz, val + (size_t)e * esz, esz);
        }
      dc.alloc(sizeof(int) * hc.size());
      dv.alloc(hv.size());
      FH_HIP(hipMemcpyAsync(dc.p, hc.data(), sizeof(int) * hc.size(), hipMemcpyHostToDevice, st));
      FH_HIP(hipMemcpyAsync(dv.p, hv.data(), hv.size(), hipMemcpyHostToDevice, st));
      FH_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
    };
    ell(m, rp, ci, (const char*)vals, M->max_row, M->csr_ell_w, M->csr_ell_ci, M->csr_ell_v);
    ell(n, cp.data(), ri.data(), cv.data(), M->max_col, M->csr_tell_w, M->csr_tell_ci, M->csr_tell_v);
  }
  FH_HIP(hipStreamSynchronize(st));
  return M;
}

DMatP make_dense(int dtype, int m, int n, const void* data, int ld, bool on_device, hipStream_t st) {
  FH_CHECK(m >= 0 && n >= 0 && ld >= std::max(m, 1), "dense: bad shape or leading dimension");
  const size_t esz = dtype_size(dtype);
  auto M = std::make_shared<DMat>();
  M->kind = K_DENSE;
  M->dtype = dtype;
  M->m = m;
  M->n = n;
  M->nnz = -1;
  M->d.alloc(esz * std::max<size_t>((size_t)m * n, 1));
  if (m > 0 && n > 0 && data)
    FH_HIP(hipMemcpy2DAsync(M->d.p, esz * m, data, esz * ld, esz * m, n,
                            on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
  FH_HIP(hipStreamSynchronize(st));
  return M;
}

// Bytes of each of the (at most two) intermediates of an FP64 chain product;
// *nbuf: how many of them the product uses.
size_t faust_ws_need_f64(const std::vector<DMatP>& f, char op, int p, int* nbuf) {
  int maxr = 0;
  for (auto& x : f) maxr = std::max(maxr, op == 'N' ? x->m : x->n);
  *nbuf = std::min(2, (int)f.size() - 1);
  return sizeof(double) * std::max<size_t>((size_t)maxr * p, 1);
}

// Grow a Faust's own intermediates to `need` bytes (hipMalloc only when a
// larger product than any before is requested) and return their addresses.
ChainWs faust_ws_grow(DBuf* ws, size_t need, int nbuf, hipStream_t st, DBuf* ctl) {
  ChainWs w{{nullptr, nullptr}, nullptr};
  for (int b = 0; b < nbuf; ++b) {
    if (ws[b].bytes < need) {
      FH_HIP(hipStreamSynchronize(st));  // the old buffer may still be in use
      ws[b].alloc(need);
    }
    w.p[b] = ws[b].p;
  }
  if (ctl) {
    if (ctl->bytes < bsr_chain_ctl_bytes()) {
      ctl->alloc(bsr_chain_ctl_bytes());
      bsr_chain_ctl_clear(ctl->as<int>(), st);
    }
    w.ctl = ctl->as<int>();
  }
  return w;
}

// Y = op(F) X for an FP64 chain; X, Y col-major with ld = rows; ws holds the
// intermediates (faust_ws_need_f64 bytes each).
// ---------------------------------------------------------------- autotune
// Execution forms of an FP64 CSR chain, all with the same bits: one launch
// per factor, the staged one-launch chain (wc columns per workgroup), the
// prefetched one-launch chain (wc columns, pf entries per row ahead).
struct ChainForm {
  int kind;  // 0 per factor, 1 staged, 2 prefetched
  int wc, pf;
};

namespace {
std::mutex g_tune_mu;
std::map<std::string, ChainForm> g_tune;  // chain signature -> quickest form measured
std::map<std::string, int> g_tune_mode;   // BSR / FP32 CSR chain signature -> FH_MUL_* mode
}  // namespace

// FH_AUTOTUNE=0 turns the measurement off (the default forms are used).
bool autotune_on() {
  const char* e = getenv("FH_AUTOTUNE");
  return !(e && e[0] == '0');
}

bool autotune_get_mode(const std::string& key, int* mode) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune_mode.find(key);
  if (it == g_tune_mode.end()) return false;
  *mode = it->second;
  return true;
}

void autotune_put_mode(const std::string& key, int mode) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_mode[key] = mode;
}

bool stream_capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  const hipError_t e = hipStreamIsCapturing(st, &cs);
  (void)hipGetLastError();
  return e != hipSuccess || cs != hipStreamCaptureStatusNone;
}

// What the quickest form depends on: direction, columns, each op'd factor's
// shape and entries, and the device.
static std::string chain_sig(char op, int p, int in_rows, const std::vector<ChainCsrOp>& ops) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::string k = std::to_string(dev) + op + ":" + std::to_string(p) + ":" + std::to_string(in_rows);
  for (const auto& o : ops) k += "," + std::to_string(o.rows) + "/" + std::to_string(o.nnz);
  return k;
}

// Y = op(F) X with the FP64 chain.  AUTO mode on an all-CSR chain uses the
// form measured quickest for this chain signature (measured on its first
// call outside graph capture: every candidate once untimed, then the minimum
// of three timings of five back-to-back products, HIP events on the launch
// stream; the default form stays unless another is at least 3% quicker);
// otherwise, and when
// measuring is off or impossible, the prefetched one-launch chain where the
// shape allows it, the staged one where the planner takes it, else one launch
// per factor.  `tune` false (allocation-free callers) never measures.
void faust_multiply_f64(const std::vector<DMatP>& f, char op, const double* X, int p, double* Y, hipStream_t st,
                        const ChainWs& ws, FaustTimers* tm, int* launches, bool tune) {
  FH_CHECK(!f.empty(), "Faust::multiply: empty Faust");
  const int J = (int)f.size();
  double* const t1 = static_cast<double*>(ws.p[0]);
  double* const t2 = static_cast<double*>(ws.p[1]);
  const int in_rows0 = (op == 'N') ? f.back()->n : f.front()->m;

  auto per_factor = [&](FaustTimers* tmr) {
    const double* in = X;
    int in_rows = in_rows0;
    for (int i = 0; i < J; ++i) {
      const DMat& S = (op == 'N') ? *f[J - 1 - i] : *f[i];
      const int out_rows = (op == 'N') ? S.m : S.n;
      FH_CHECK(((op == 'N') ? S.n : S.m) == in_rows, "Faust::multiply: factor dimensions do not chain");
      double* out = (i == J - 1) ? Y : ((i & 1) ? t2 : t1);
      FH_CHECK(S.dtype == FH_F64, "Faust::multiply: FP64 chain expected");
      const size_t slot = (op == 'N') ? (size_t)(J - 1 - i) : (size_t)i;
      if (tmr) tmr->begin(slot, st);
      if (S.kind == K_CSR)
        csr_spmm_f64(S, op, in, in_rows, p, out, out_rows, 1.0, 0.0, st);
      else if (S.kind == K_DENSE)
        gemm_f64(op == 'N' ? 'N' : 'T', 'N', out_rows, p, in_rows, 1.0, S.d.as<double>(), S.m, in, in_rows, 0.0,
                 out, out_rows, st);
      else
        throw Error(FH_ERR_UNSUPPORTED, "Faust::multiply: BSR factors need the FP32/BF16 chain");
      if (tmr) tmr->end(slot, st);
      in = out;
      in_rows = out_rows;
    }
  };

  // the chain as op'd CSR factors in application order, when it is one
  std::vector<ChainCsrOp> ops;
  int r = in_rows0, widest = 0, nnz_max = 0;
  double mean_max = 0.0;
  bool ok = true;
  {
    const char* e_ell = getenv("FH_CSR_ELL");
    const bool ell_on = !(e_ell && e_ell[0] == '0');
    for (int i = 0; i < J && ok; ++i) {
      const DMat& S = (op == 'N') ? *f[J - 1 - i] : *f[i];
      ok = S.kind == K_CSR && S.dtype == FH_F64 && ((op == 'N') ? S.n : S.m) == r && (op == 'N' || S.has_csc);
      if (!ok) break;
      const bool tr = op != 'N';
      const int wr = tr ? S.max_col : S.max_row;
      widest = std::max(widest, wr < 0 ? (1 << 30) : wr);
      const int rows = tr ? S.n : S.m;
      mean_max = std::max(mean_max, (double)S.nnz / std::max(rows, 1));
      nnz_max = std::max(nnz_max, (int)S.nnz);
      const int ew = tr ? S.csr_tell_w : S.csr_ell_w;
      ops.push_back({tr ? S.cp.as<int>() : S.rp.as<int>(), tr ? S.ri.as<int>() : S.ci.as<int>(),
                     tr ? S.cv.as<double>() : S.v.as<double>(), rows, (int)S.nnz,
                     ell_on && ew > 0 && ew % 8 == 0 && ew <= 32 ? 1 : 0});
      r = rows;
    }
  }
  const int out_rows = r;
  auto run = [&](const ChainForm& c, FaustTimers* tmr) {
    if (c.kind == 0) {
      per_factor(tmr);
      return;
    }
    if (tmr) tmr->begin(0, st);
    if (c.kind == 1)
      csr_chain_run_f64(ops, in_rows0, c.wc, X, in_rows0, p, Y, out_rows, 1.0, 0.0, st);
    else
      csr_chain_pf_run_f64(ops, in_rows0, c.wc, c.pf, X, in_rows0, p, Y, out_rows, 1.0, 0.0, st);
    if (tmr) tmr->end(0, st);
  };
  auto done = [&](const ChainForm& c) {
    if (launches) *launches = c.kind == 0 ? J : 1;
  };

  if (ok) {
    // default form: prefetched (8 entries ahead unless the rows are long on
    // average and some exceed 8), else the staged plan, else per factor
    ChainForm def{0, 0, 0};
    const int pwc = csr_chain_pf_plan_f64(ops, in_rows0, p);
    const int swc = pwc > 0 ? 0 : csr_chain_plan_f64(ops, in_rows0, p);
    if (pwc > 0)
      def = {2, pwc, (widest <= 8 || mean_max <= 4.0) ? 8 : 16};
    else if (swc > 0)
      def = {1, swc, 0};
    ChainForm use = def;
    const bool tunable = mul_mode() == FH_MUL_AUTO && autotune_on() && !getenv("FH_CHAIN_FUSED") &&
                         !(getenv("FH_CHAIN_PF") && getenv("FH_CHAIN_PF")[0] == '0');
    if (tunable) {
      const std::string key = chain_sig(op, p, in_rows0, ops);
      bool known = false;
      {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
          use = it->second;
          known = true;
        }
      }
      if (!known && tune && !stream_capturing(st)) {
        // the default form first: another replaces the quickest so far only
        // when it is at least 3% quicker, so near-ties keep the default and
        // the choice does not flip between processes on timing noise
        std::vector<ChainForm> cand{def};
        auto add = [&](ChainForm c) {
          if (c.kind != def.kind || c.wc != def.wc || c.pf != def.pf) cand.push_back(c);
        };
        add({0, 0, 0});
        const int fwc = csr_chain_fit_wc(ops, in_rows0, p);
        if (fwc > 0) add({1, fwc, 0});
        if (csr_chain_pf_fits(ops, in_rows0, p))
          for (int pf : {8, 16}) {
            if (pf == 16 && widest <= 8) continue;
            for (int wc : {1, 2, 4})
              if (wc == 1 || (p + wc - 1) / wc >= 8) add({2, wc, pf});
          }
        hipEvent_t e0, e1;
        FH_HIP(hipEventCreate(&e0));
        FH_HIP(hipEventCreate(&e1));
        float best_ms = 1e30f;
        // each timing covers five back-to-back products: the steady state a
        // caller's loop or graph sees
        for (const ChainForm& c : cand) {
          run(c, nullptr);
          float m = 1e30f;
          for (int rep = 0; rep < 3; ++rep) {
            FH_HIP(hipEventRecord(e0, st));
            for (int r = 0; r < 5; ++r) run(c, nullptr);
            FH_HIP(hipEventRecord(e1, st));
            FH_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            FH_HIP(hipEventElapsedTime(&ms, e0, e1));
            m = std::min(m, ms);
          }
          if (m < 0.97f * best_ms) {
            best_ms = m;
            use = c;
          }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[key] = use;
      }
    }
    run(use, tm);
    done(use);
    return;
  }
  per_factor(tm);
  if (launches) *launches = J;
}

void autotune_clear() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
  g_tune_mode.clear();
}

// Transform::get_product (faust_Transform.hpp:599-626)
DMatP faust_get_product_f64(const std::vector<DMatP>& f, hipStream_t st) {
  FH_CHECK(!f.empty(), "get_product: empty Faust");
  const size_t n = f.size();
  if (n == 1) return dense_of(*f[0], st);
  auto mat = std::make_shared<DMat>();
  matmul_f64(*f[n - 2], 'N', *f[n - 1], 'N', 1.0, 0.0, *mat, st);
  for (int i = (int)n - 3; i >= 0; --i) {
    auto nm = std::make_shared<DMat>();
    matmul_f64(*f[i], 'N', *mat, 'N', 1.0, 0.0, *nm, st);
    mat = nm;
  }
  return mat;
}

static bool chain_is_zero(const std::vector<DMatP>& f, hipStream_t st) {
  for (auto& x : f)
    if (nnz_of(*x, st) == 0) return true;
  return false;
}

// the power-iteration chain of F^H F (or F F^H, the smaller side); false when
// the norm is known without it (*known: 1 for an empty Faust, 0 for a zero one)
static bool spectral_chain(const std::vector<DMatP>& f, std::vector<ChainOp>& ops, int& n, double* known,
                           hipStream_t st) {
  if (f.empty()) {
    *known = 1.0;
    return false;
  }
  if (chain_is_zero(f, st)) {
    *known = 0.0;
    return false;
  }
  const int nr = f.front()->m, nc = f.back()->n;
  if (nc < nr) {
    for (int i = (int)f.size() - 1; i >= 0; --i) ops.push_back({f[i].get(), 'T'});
    for (auto& x : f) ops.push_back({x.get(), 'N'});
    n = nc;
  } else {
    for (auto& x : f) ops.push_back({x.get(), 'N'});
    for (int i = (int)f.size() - 1; i >= 0; --i) ops.push_back({f[i].get(), 'T'});
    n = nr;
  }
  return true;
}

// Transform::spectralNorm (faust_Transform.hpp:684-709)
double faust_spectral_norm_f64(const std::vector<DMatP>& f, int maxit, double thr, int* flag, hipStream_t st) {
  std::vector<ChainOp> ops;
  int n = 0;
  double known;
  if (!spectral_chain(f, ops, n, &known, st)) return known;
  double lam = power_iteration_f64(ops, n, maxit, thr, flag, st);
  return std::fabs(std::sqrt(lam));
}

// the 2-norms of two Fausts (PALM's R and L of one factor update), their
// power iterations sharing a launch; the same values as two calls of
// faust_spectral_norm_f64 (f1 first), flag from the last one computed
void faust_spectral_norm_pair_f64(const std::vector<DMatP>& f1, const std::vector<DMatP>& f2, int maxit, double thr,
                                  int* flag, double* n1, double* n2, hipStream_t st) {
  std::vector<ChainOp> o1, o2;
  int m1 = 0, m2 = 0;
  double k1, k2;
  const bool p1 = spectral_chain(f1, o1, m1, &k1, st), p2 = spectral_chain(f2, o2, m2, &k2, st);
  if (p1 && p2) {
    double l1, l2;
    power_iteration_pair_f64(o1, m1, o2, m2, maxit, thr, &l1, &l2, flag, st);
    *n1 = std::fabs(std::sqrt(l1));
    *n2 = std::fabs(std::sqrt(l2));
    return;
  }
  *n1 = p1 ? std::fabs(std::sqrt(power_iteration_f64(o1, m1, maxit, thr, flag, st))) : k1;
  *n2 = p2 ? std::fabs(std::sqrt(power_iteration_f64(o2, m2, maxit, thr, flag, st))) : k2;
}

}  // namespace fh

namespace fh {
static thread_local int tl_mul_mode = FH_MUL_AUTO;
int mul_mode() { return tl_mul_mode; }
MulModeScope::MulModeScope(int m) : prev(tl_mul_mode) { tl_mul_mode = m; }
MulModeScope::~MulModeScope() { tl_mul_mode = prev; }
}  // namespace fh

extern "C" {

int fh_mat_dense_create(int dtype, int nrows, int ncols, const void* data, int ld, int src_on_device, fh_mat_t* out) {
  FH_API_BEGIN
  FH_CHECK(out, "null output handle");
  FH_CHECK(dtype == FH_F64 || dtype == FH_F32 || dtype == FH_BF16, "unknown dtype");
  auto M = make_dense(dtype, nrows, ncols, data, ld, src_on_device != 0, nullptr);
  *out = new fh_mat_s{M};
  FH_API_END
}

int fh_mat_csr_create(int dtype, int nrows, int ncols, int nnz, const int32_t* rowptr, const int32_t* colind,
                      const void* vals, int src_on_device, fh_mat_t* out) {
  FH_API_BEGIN
  FH_CHECK(out && rowptr && (nnz == 0 || (colind && vals)), "null argument");
  FH_CHECK(dtype == FH_F64 || dtype == FH_F32 || dtype == FH_BF16, "unknown dtype");
  const size_t esz = dtype_size(dtype);
  if (src_on_device) {
    std::vector<int32_t> rp(nrows + 1), ci(nnz);
    std::vector<char> v(esz * nnz);
    FH_HIP(hipMemcpy(rp.data(), rowptr, sizeof(int) * (nrows + 1), hipMemcpyDeviceToHost));
    if (nnz) {
      FH_HIP(hipMemcpy(ci.data(), colind, sizeof(int) * nnz, hipMemcpyDeviceToHost));
      FH_HIP(hipMemcpy(v.data(), vals, esz * nnz, hipMemcpyDeviceToHost));
    }
    *out = new fh_mat_s{make_csr_from_host(dtype, nrows, ncols, nnz, rp.data(), ci.data(), v.data(), nullptr)};
  } else {
    *out = new fh_mat_s{make_csr_from_host(dtype, nrows, ncols, nnz, rowptr, colind, vals, nullptr)};
  }
  FH_API_END
}

int fh_mat_info(fh_mat_t M, int* kind, int* dtype, int* nrows, int* ncols, int64_t* nnz) {
  FH_API_BEGIN
  FH_CHECK(M, "null handle");
  const DMat& d = *M->m;
  if (kind) *kind = (int)d.kind;
  if (dtype) *dtype = d.dtype;
  if (nrows) *nrows = d.m;
  if (ncols) *ncols = d.n;
  if (nnz) {
    if (d.kind == K_DENSE)
      *nnz = nnz_of(*M->m, nullptr);
    else if (d.kind == K_BSR)
      *nnz = (int64_t)d.bnnz * d.bm * d.bn;
    else
      *nnz = d.nnz;
  }
  FH_API_END
}

int fh_mat_tocpu(fh_mat_t M, void* data_or_vals, int32_t* rowptr, int32_t* colind) {
  FH_API_BEGIN
  FH_CHECK(M, "null handle");
  const DMat& d = *M->m;
  const size_t esz = dtype_size(d.dtype);
  if (d.kind == K_DENSE) {
    if (d.m * (size_t)d.n)
      FH_HIP(hipMemcpy(data_or_vals, d.d.p, esz * (size_t)d.m * d.n, hipMemcpyDeviceToHost));
  } else if (d.kind == K_CSR) {
    if (rowptr) FH_HIP(hipMemcpy(rowptr, d.rp.p, sizeof(int) * (d.m + 1), hipMemcpyDeviceToHost));
    if (d.nnz) {
      if (colind) FH_HIP(hipMemcpy(colind, d.ci.p, sizeof(int) * d.nnz, hipMemcpyDeviceToHost));
      if (data_or_vals) FH_HIP(hipMemcpy(data_or_vals, d.v.p, esz * d.nnz, hipMemcpyDeviceToHost));
    }
  } else {
    bsr_to_host(d, data_or_vals, rowptr, colind);
  }
  FH_API_END
}

int fh_mat_to_dense(fh_mat_t M, void* dense_dev, void* stream) {
  FH_API_BEGIN
  FH_CHECK(M && dense_dev, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const DMat& d = *M->m;
  if (d.dtype == FH_F64) {
    if (d.kind == K_DENSE)
      FH_HIP(hipMemcpyAsync(dense_dev, d.d.p, sizeof(double) * (size_t)d.m * d.n, hipMemcpyDeviceToDevice, st));
    else if (d.kind == K_CSR)
      csr_to_dense_f64(d, (double*)dense_dev, st);
    else
      throw Error(FH_ERR_UNSUPPORTED, "to_dense: FP64 BSR not supported");
  } else {
    mixed_to_dense(d, dense_dev, st);
  }
  FH_API_END
}

int fh_mat_free(fh_mat_t M) {
  FH_API_BEGIN
  delete M;
  FH_API_END
}

int fh_mat_mul_dense(fh_mat_t M, char op, const void* X, int ldx, int ncols, void* Y, int ldy, const void* alpha,
                     const void* beta, void* stream) {
  FH_API_BEGIN
  FH_CHECK(M, "null argument");
  FH_CHECK(op == 'N' || op == 'T' || op == 'H', "op must be N, T or H");
  FH_CHECK(ncols >= 0, "negative column count");
  if (ncols == 0) return FH_OK;  // nothing to compute (X and Y may be null)
  FH_CHECK(X && Y, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const DMat& S = *M->m;
  if (S.dtype == FH_F64) {
    const double a = alpha ? *(const double*)alpha : 1.0, b = beta ? *(const double*)beta : 0.0;
    const int orows = op == 'N' ? S.m : S.n, irows = op == 'N' ? S.n : S.m;
    FH_CHECK(ldx >= std::max(irows, 1) && ldy >= std::max(orows, 1), "bad leading dimension");
    if (S.kind == K_CSR)
      csr_spmm_f64(S, op, (const double*)X, ldx, ncols, (double*)Y, ldy, a, b, st);
    else if (S.kind == K_DENSE)
      gemm_f64(op == 'N' ? 'N' : 'T', 'N', orows, ncols, irows, a, S.d.as<double>(), S.m, (const double*)X, ldx, b,
               (double*)Y, ldy, st);
    else
      throw Error(FH_ERR_UNSUPPORTED, "FP64 BSR products are not supported");
  } else {
    mixed_mul_dense(S, op, X, ldx, ncols, Y, ldy, alpha, beta, st);
  }
  FH_API_END
}

int fh_gemm(int dtype, char opA, char opB, int m, int n, int k, const void* alpha, const void* A, int lda,
            const void* B, int ldb, const void* beta, void* C, int ldc, void* stream) {
  FH_API_BEGIN
  FH_CHECK(dtype == FH_F64, "fh_gemm: FP64 only");
  FH_CHECK(m >= 0 && n >= 0 && k >= 0, "fh_gemm: negative size");
  gemm_f64(opA, opB, m, n, k, alpha ? *(const double*)alpha : 1.0, (const double*)A, lda, (const double*)B, ldb,
           beta ? *(const double*)beta : 0.0, (double*)C, ldc, (hipStream_t)stream);
  FH_API_END
}

int fh_faust_create(fh_faust_t* out) {
  FH_API_BEGIN
  FH_CHECK(out, "null output handle");
  *out = new fh_faust_s();
  FH_API_END
}

int fh_faust_push_back(fh_faust_t F, fh_mat_t M) {
  FH_API_BEGIN
  FH_CHECK(F && M, "null handle");
  if (!F->f.empty()) FH_CHECK(F->f.back()->n == M->m->m, "push_back: dimensions do not chain");
  F->f.push_back(M->m);
  FH_API_END
}

int fh_faust_size(fh_faust_t F, int* n) {
  FH_API_BEGIN
  FH_CHECK(F && n, "null argument");
  *n = (int)F->f.size();
  FH_API_END
}

int fh_faust_factor(fh_faust_t F, int i, fh_mat_t* out) {
  FH_API_BEGIN
  FH_CHECK(F && out, "null argument");
  FH_CHECK(i >= 0 && i < (int)F->f.size(), "factor index out of range");
  auto* h = new fh_mat_s();
  h->m = F->f[i];
  *out = h;
  FH_API_END
}

int fh_faust_free(fh_faust_t F) {
  FH_API_BEGIN
  delete F;
  FH_API_END
}

int fh_faust_set_mul_mode(fh_faust_t F, int mode) {
  FH_API_BEGIN
  FH_CHECK(F, "null handle");
  const bool ok = mode == FH_MUL_AUTO || mode == FH_MUL_PER_FACTOR || mode == FH_MUL_FUSED ||
                  (mode >= FH_MUL_BSR_KERNEL(0) && mode <= FH_MUL_BSR_KERNEL(5)) || mode == FH_MUL_CSR_ROWS1 ||
                  mode == FH_MUL_CSR_ROWS2;
  FH_CHECK(ok, "fh_faust_set_mul_mode: unknown mode");
  F->mul_mode = mode;
  FH_API_END
}

int fh_faust_set_act_dtype(fh_faust_t F, int dtype) {
  FH_API_BEGIN
  FH_CHECK(F, "null handle");
  FH_CHECK(dtype == FH_BF16 || dtype == FH_F32, "activation dtype must be FH_BF16 or FH_F32");
  if (dtype == FH_F32)
    for (auto& x : F->f)
      if (x->kind != K_BSR || x->dtype != FH_F32)
        throw Error(FH_ERR_UNSUPPORTED, "FP32 activations are offered for chains of FP32 BSR factors");
  for (auto& x : F->f) FH_CHECK(x->kind == K_BSR, "the activation dtype is chosen for BSR chains");
  F->act_dtype = dtype;
  FH_API_END
}

int fh_faust_get_act_dtype(fh_faust_t F, int* dtype) {
  FH_API_BEGIN
  FH_CHECK(F && dtype, "null argument");
  FH_CHECK(!F->f.empty(), "empty Faust");
  if (F->f.front()->dtype == FH_F64)
    *dtype = FH_F64;
  else if (F->f.front()->kind == K_BSR)
    *dtype = F->act_dtype == FH_F32 ? FH_F32 : FH_BF16;
  else
    *dtype = FH_F32;
  FH_API_END
}

int fh_faust_get_mul_mode(fh_faust_t F, int* mode) {
  FH_API_BEGIN
  FH_CHECK(F && mode, "null argument");
  *mode = F->mul_mode;
  FH_API_END
}

// A dense FP32 factor in an FP32 chain runs as its CSR twin (built once from
// a host copy of the factor, kept with it): the FP32 chain kernels take CSR.
static std::vector<DMatP> fp32_chain_factors(const std::vector<DMatP>& f, hipStream_t st) {
  std::vector<DMatP> out = f;
  for (auto& x : out) {
    if (x->kind != K_DENSE || x->dtype != FH_F32) continue;
    if (!x->csr_twin) {
      const int m = x->m, n = x->n;
      std::vector<float> h((size_t)m * n);
      if (!h.empty()) FH_HIP(hipMemcpyAsync(h.data(), x->d.p, sizeof(float) * h.size(), hipMemcpyDeviceToHost, st));
      FH_HIP(hipStreamSynchronize(st));
      std::vector<int32_t> rp(m + 1, 0), ci;
      std::vector<float> v;
      for (int r = 0; r < m; ++r) {
        for (int c = 0; c < n; ++c) {
          const float a = h[(size_t)c * m + r];
          if (a != 0.0f) {
            ci.push_back(c);
            v.push_back(a);
          }
        }
        rp[r + 1] = (int32_t)ci.size();
      }
      x->csr_twin = make_csr_from_host(FH_F32, m, n, (int)ci.size(), rp.data(), ci.data(), v.data(), st);
    }
    x = x->csr_twin;
  }
  return out;
}

int fh_faust_multiply(fh_faust_t F, char op, const void* X, int p, void* Y, void* stream) {
  FH_API_BEGIN
  FH_CHECK(F, "null argument");
  MulModeScope mode_scope(F->mul_mode);
  FH_CHECK(op == 'N' || op == 'T' || op == 'H', "op must be N, T or H");
  FH_CHECK(!F->f.empty(), "Faust::multiply: empty Faust");
  FH_CHECK(p >= 0, "negative column count");
  if (p == 0) {  // nothing to compute (X and Y may be null)
    F->last_launches = 0;
    return FH_OK;
  }
  FH_CHECK(X && Y, "null argument");
  hipStream_t st = (hipStream_t)stream;
  int nbuf = 0;
  if (F->f.front()->dtype == FH_F64) {
    const size_t need = faust_ws_need_f64(F->f, op, p, &nbuf);
    const ChainWs ws = faust_ws_grow(F->ws, need, nbuf, st);
    faust_multiply_f64(F->f, op, (const double*)X, p, (double*)Y, st, ws, &F->timers, &F->last_launches);
  } else {
    const std::vector<DMatP> f = fp32_chain_factors(F->f, st);
    const bool act_f32 = F->act_dtype == FH_F32;
    const size_t need = faust_ws_need_mixed(f, op, p, &nbuf, act_f32);
    const ChainWs ws =
        faust_ws_grow(F->ws, need, nbuf, st, (f.front()->kind == K_BSR && !act_f32) ? &F->ctl : nullptr);
    F->last_launches = faust_multiply_mixed(f, op, X, p, Y, st, ws, &F->timers, act_f32);
  }
  FH_API_END
}

int fh_faust_last_launches(fh_faust_t F, int* launches) {
  FH_API_BEGIN
  FH_CHECK(F && launches, "null argument");
  *launches = F->last_launches;
  FH_API_END
}

int fh_faust_enable_timers(fh_faust_t F, int on) {
  FH_API_BEGIN
  FH_CHECK(F, "null handle");
  if (!on) F->timers.collect();
  F->timers.on = on != 0;
  F->timers.resize(F->f.size() + 1);
  FH_API_END
}

int fh_faust_timers(fh_faust_t F, double* ms, int64_t* counts) {
  FH_API_BEGIN
  FH_CHECK(F && ms && counts, "null argument");
  F->timers.resize(F->f.size() + 1);
  F->timers.collect();
  for (size_t s = 0; s <= F->f.size(); ++s) {
    ms[s] = F->timers.ms[s];
    counts[s] = F->timers.count[s];
    F->timers.ms[s] = 0.0;
    F->timers.count[s] = 0;
  }
  FH_API_END
}

int fh_faust_get_product(fh_faust_t F, void* out_dev, void* stream) {
  FH_API_BEGIN
  FH_CHECK(F && out_dev, "null argument");
  for (auto& x : F->f)
    if (x->dtype != FH_F64 || x->kind == K_BSR)
      throw Error(FH_ERR_UNSUPPORTED, "get_product: FP64 dense / CSR chains only (multiply other chains by the identity)");
  hipStream_t st = (hipStream_t)stream;
  DMatP P = faust_get_product_f64(F->f, st);
  FH_HIP(hipMemcpyAsync(out_dev, P->d.p, sizeof(double) * (size_t)P->m * P->n, hipMemcpyDeviceToDevice, st));
  FH_HIP(hipStreamSynchronize(st));
  FH_API_END
}

int fh_faust_spectral_norm(fh_faust_t F, int max_iter, double threshold, double* norm, int* flag, void* stream) {
  FH_API_BEGIN
  FH_CHECK(F && norm, "null argument");
  for (auto& x : F->f)
    if (x->dtype != FH_F64 || x->kind == K_BSR)
      throw Error(FH_ERR_UNSUPPORTED, "spectral_norm: FP64 dense / CSR chains only");
  *norm = faust_spectral_norm_f64(F->f, max_iter, threshold, flag, (hipStream_t)stream);
  FH_API_END
}

int fh_prox(int dtype, int name, void* M, int nrows, int ncols, int64_t k, int normalized, int pos, const void* supp,
            void* stream) {
  FH_API_BEGIN
  FH_CHECK(dtype == FH_F64, "fh_prox: FP64 only");
  FH_CHECK(M && nrows >= 0 && ncols >= 0, "bad argument");
  prox_f64(name, (double*)M, nrows, ncols, k, normalized != 0, pos != 0, (const double*)supp, (hipStream_t)stream);
  FH_API_END
}

int fh_prox_real(int dtype, int name, void* M, int nrows, int ncols, double s, int normalized, int pos,
                 void* stream) {
  FH_API_BEGIN
  FH_CHECK(dtype == FH_F64, "fh_prox_real: FP64 only");
  FH_CHECK(M && nrows >= 0 && ncols >= 0, "bad argument");
  prox_real_f64(name, (double*)M, nrows, ncols, s, normalized != 0, pos != 0, (hipStream_t)stream);
  FH_API_END
}

int fh_norm_frob(int dtype, const void* M, int64_t n, double* out, void* stream) {
  FH_API_BEGIN
  FH_CHECK(dtype == FH_F64 && out, "fh_norm_frob: FP64 only");
  hipStream_t st = (hipStream_t)stream;
  DBuf s(sizeof(double));
  reduce_dot_f64((const double*)M, (const double*)M, n, s.as<double>(), st);
  double h;
  FH_HIP(hipMemcpyAsync(&h, s.p, sizeof(double), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  *out = std::sqrt(h);
  FH_API_END
}

int fh_dot(int dtype, const void* x, const void* y, int64_t n, double* out, void* stream) {
  FH_API_BEGIN
  FH_CHECK(dtype == FH_F64 && out, "fh_dot: FP64 only");
  hipStream_t st = (hipStream_t)stream;
  DBuf s(sizeof(double));
  reduce_dot_f64((const double*)x, (const double*)y, n, s.as<double>(), st);
  FH_HIP(hipMemcpyAsync(out, s.p, sizeof(double), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  FH_API_END
}

}  // extern "C"
