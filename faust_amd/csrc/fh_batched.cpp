// fh_batched.cpp -- host driver of the batched PALM4MSA engine.
//
// Builds the per-iteration schedule of batched kernels once (left-to-right
// updates with packed partial products, as palm4msa2 with the
// ParamsPalm4MSA defaults: faust_palm4msa2020.hpp:47-300), captures it into
// a HIP graph on a private stream, and replays the graph for every
// iteration.  Tile sizes of the grouped GEMM are picked per target size by a
// timing autotuner at set-up.  The mixed-precision scheme runs the main
// iterations in FP32, then refinement sweeps in FP64 with every factor
// projected on its frozen FP32 support (prox supp), and verifies the final
// relative error in FP64.
#include <algorithm>
#include <chrono>
#include <map>
#include <numeric>
#include <type_traits>

#include "fh_api_common.hpp"
#include "fh_batched.hpp"

namespace fh {

struct BatchSpec {
  int nt, nf;
  std::vector<int> n;
};

template <class T>
struct Batch {
  BatchSpec sp;
  int slots;                       // matrices per target
  std::vector<int64_t> off;        // element offset of each target's block
  DBuf arena, lam, nR, nL, relerr, powv;
  std::vector<int64_t> powoff;     // per target start of its 2*nf power vectors
  // schedule
  struct Stage {
    int kind;  // 0 gemm, 1 power(R->nR), 2 power(L->nL), 3 prox, 4 lambda, 5 sparse build, 6 / 7 sparse power (R / L),
               // 8 sparse x dense product
    DBuf jobs;
    int njobs = 0;
    DBuf tiles;  // kind 8: (job, first column, columns) per workgroup
    int ntiles = 0;
    int nmax = 0;  // kinds 6 / 7: the largest target edge of the jobs
    int mc = 0;    // kinds 6 / 7: the factor count all jobs with factors share (1 or 2, edges <= 512), else 0
  };
  std::vector<Stage> stages;
  DBuf tiles32, tiles64, tilesM;
  int nt32 = 0, nt64 = 0, ntM = 0;
  std::vector<int> tile_of;  // per target: 32 or 64 (VALU), BGEMM_MFMA (FP32 MFMA, 64 x 64)
  int prox_mode = 0;
  int maxit = 100;
  double thr = 1e-6;
  int warm = 1;
  // step-size power iterations on the sparse factors instead of the dense
  // partial products (FP32 sp batches): every factor's compressed rows and
  // columns, rebuilt after its projection
  bool sparse = false;
  DBuf spmem, spfac;  // the arrays; the BSpFac<T> of target b, factor f at b * nf + f
  std::vector<BSpFac<T>> spfac_h;

  T* mat(int b, int slot) const { return arena.as<T>() + off[b] + (int64_t)slot * sp.n[b] * sp.n[b]; }
  // slot layout: 0 A, 1..nf S, nf+1..2nf-1 PR[0..nf-2], 2nf..3nf-2 PL[1..nf-1], then DR, TMP, TR, SM, MASK[nf]
  int sA() const { return 0; }
  int sS(int f) const { return 1 + f; }
  int sPR(int i) const { return sp.nf + 1 + i; }
  int sPL(int f) const { return 2 * sp.nf + f - 1; }
  int sDR() const { return 3 * sp.nf - 1; }
  int sTMP() const { return 3 * sp.nf; }
  int sTR() const { return 3 * sp.nf + 1; }
  int sSM() const { return 3 * sp.nf + 2; }
  int sMASK(int f) const { return 3 * sp.nf + 3 + f; }
};

template <class T>
static void batch_alloc(Batch<T>& B, const BatchSpec& sp, bool masks) {
  B.sp = sp;
  B.slots = 3 * sp.nf + 3 + (masks ? sp.nf : 0);
  B.off.resize(sp.nt);
  int64_t tot = 0;
  for (int b = 0; b < sp.nt; ++b) {
    FH_CHECK(sp.n[b] > 0 && sp.n[b] <= BATCH_NMAX, "batched PALM: target size out of range (1..2048)");
    B.off[b] = tot;
    tot += (int64_t)B.slots * sp.n[b] * sp.n[b];
  }
  B.arena.alloc(sizeof(T) * std::max<int64_t>(tot, 1));
  FH_HIP(hipMemset(B.arena.p, 0, sizeof(T) * std::max<int64_t>(tot, 1)));
  // the engine works on a non-blocking stream that does not wait for the
  // null stream: the clear must be complete before any of its work starts
  FH_HIP(hipDeviceSynchronize());
  B.lam.alloc(sizeof(double) * sp.nt);
  B.nR.alloc(sizeof(double) * sp.nt);
  B.nL.alloc(sizeof(double) * sp.nt);
  B.relerr.alloc(sizeof(double) * sp.nt);
  B.powoff.resize(sp.nt);
  int64_t pt = 0;
  for (int b = 0; b < sp.nt; ++b) {
    B.powoff[b] = pt;
    pt += (int64_t)2 * sp.nf * sp.n[b];
  }
  B.powv.alloc(sizeof(double) * std::max<int64_t>(pt, 1));
  std::vector<double> hv(std::max<int64_t>(pt, 1));
  for (int b = 0; b < sp.nt; ++b) {
    const std::vector<double>& x0 = power_x0(sp.n[b]);
    for (int s = 0; s < 2 * sp.nf; ++s)
      std::copy(x0.begin(), x0.end(), hv.begin() + B.powoff[b] + (int64_t)s * sp.n[b]);
  }
  FH_HIP(hipMemcpy(B.powv.p, hv.data(), sizeof(double) * hv.size(), hipMemcpyHostToDevice));
}

template <class T>
static void build_tiles(Batch<T>& B) {
  std::vector<int3> t32, t64, tm;
  for (int b = 0; b < B.sp.nt; ++b) {
    const int code = B.tile_of[b], n = B.sp.n[b];
    const int tile = code == 32 ? 32 : 64;
    const int nt = (n + tile - 1) / tile;
    std::vector<int3>& l = code == 32 ? t32 : (code == 64 ? t64 : tm);
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j) l.push_back(make_int3(b, i, j));
  }
  auto up = [](DBuf& d, int& cnt, const std::vector<int3>& v) {
    cnt = (int)v.size();
    d.alloc(sizeof(int3) * std::max<size_t>(v.size(), 1));
    if (!v.empty()) FH_HIP(hipMemcpy(d.p, v.data(), sizeof(int3) * v.size(), hipMemcpyHostToDevice));
  };
  up(B.tiles32, B.nt32, t32);
  up(B.tiles64, B.nt64, t64);
  up(B.tilesM, B.ntM, tm);
}

// The factor whose dense storage p is (of target b), or -1.
template <class T>
static int factor_of(const Batch<T>& B, int b, const T* p) {
  for (int f = 0; f < B.sp.nf; ++f)
    if (p == B.mat(b, B.sS(f))) return f;
  return -1;
}

template <class T>
static void add_gemm(Batch<T>& B, std::function<BGemmJob<T>(int)> mk) {
  std::vector<BGemmJob<T>> h(B.sp.nt);
  for (int b = 0; b < B.sp.nt; ++b) h[b] = mk(b);
  // a product with a factor as an operand runs through the factor's
  // compressed form when the batch keeps them (their structure is the
  // factor's value at that point of the sweep: a factor is only read before
  // its gradient step or after its projection and rebuild)
  if (B.sparse) {
    const int fa = factor_of(B, 0, h[0].A), fb = factor_of(B, 0, h[0].B);
    const int side = fa >= 0 ? 0 : (fb >= 0 ? 1 : -1);
    if (side >= 0) {
      std::vector<BSpmmJob<T>> hs(B.sp.nt);
      std::vector<int3> tl;
      for (int b = 0; b < B.sp.nt; ++b) {
        const BGemmJob<T>& g = h[b];
        const int f = side == 0 ? factor_of(B, b, g.A) : factor_of(B, b, g.B);
        FH_CHECK(f >= 0, "batched PALM: inconsistent sparse operand");
        hs[b] = BSpmmJob<T>{B.spfac_h[(size_t)b * B.sp.nf + f], side == 0 ? g.B : g.A, g.Cin, g.Cout, B.sp.n[b],
                            side, side == 0 ? g.transA : g.transB, g.alpha_mode, g.beta_mode, g.target};
        // transposed dense operands do not occur next to a factor in the schedule
        FH_CHECK((side == 0 ? g.transB : g.transA) == 0, "batched PALM: transposed dense operand of a sparse product");
        // about 8 K output entries per workgroup (the staged factor serves all of them)
        const int n = B.sp.n[b], nc = std::max(1, std::min(n, 8192 / n));
        for (int j0 = 0; j0 < n; j0 += nc) tl.push_back(make_int3(b, j0, std::min(nc, n - j0)));
      }
      typename Batch<T>::Stage s;
      s.kind = 8;
      s.njobs = B.sp.nt;
      s.jobs.alloc(sizeof(BSpmmJob<T>) * hs.size());
      FH_HIP(hipMemcpy(s.jobs.p, hs.data(), sizeof(BSpmmJob<T>) * hs.size(), hipMemcpyHostToDevice));
      s.ntiles = (int)tl.size();
      s.tiles.alloc(sizeof(int3) * tl.size());
      FH_HIP(hipMemcpy(s.tiles.p, tl.data(), sizeof(int3) * tl.size(), hipMemcpyHostToDevice));
      B.stages.push_back(std::move(s));
      return;
    }
  }
  typename Batch<T>::Stage s;
  s.kind = 0;
  s.njobs = B.sp.nt;
  s.jobs.alloc(sizeof(BGemmJob<T>) * h.size());
  FH_HIP(hipMemcpy(s.jobs.p, h.data(), sizeof(BGemmJob<T>) * h.size(), hipMemcpyHostToDevice));
  B.stages.push_back(std::move(s));
}

template <class T>
static void add_power(Batch<T>& B, int kind, std::function<const T*(int)> P, int slot) {
  std::vector<BPowJob<T>> h(B.sp.nt);
  double* out = (kind == 1 ? B.nR : B.nL).template as<double>();
  for (int b = 0; b < B.sp.nt; ++b)
    h[b] = BPowJob<T>{P(b), B.sp.n[b], b, out, B.powv.template as<double>() + B.powoff[b] + (int64_t)slot * B.sp.n[b]};
  typename Batch<T>::Stage s;
  s.kind = kind;
  s.njobs = B.sp.nt;
  s.jobs.alloc(sizeof(BPowJob<T>) * h.size());
  FH_HIP(hipMemcpy(s.jobs.p, h.data(), sizeof(BPowJob<T>) * h.size(), hipMemcpyHostToDevice));
  B.stages.push_back(std::move(s));
}

template <class T>
static void add_prox(Batch<T>& B, int f, const std::vector<int64_t>& k) {
  std::vector<BProxJob<T>> h(B.sp.nt);
  for (int b = 0; b < B.sp.nt; ++b) {
    const int64_t len = (int64_t)B.sp.n[b] * B.sp.n[b];
    h[b] = BProxJob<T>{B.mat(b, B.sS(f)), B.prox_mode == 1 ? B.mat(b, B.sMASK(f)) : nullptr, len,
                       k.empty() ? len : k[b], nullptr, nullptr, 0};
    if (B.sparse) {
      const BSpFac<T>& F = B.spfac_h[(size_t)b * B.sp.nf + f];
      h[b].klist = F.klist;
      h[b].kcount = F.kcount;
      h[b].kcap = F.cap;
    }
  }
  typename Batch<T>::Stage s;
  s.kind = 3;
  s.njobs = B.sp.nt;
  s.jobs.alloc(sizeof(BProxJob<T>) * h.size());
  FH_HIP(hipMemcpy(s.jobs.p, h.data(), sizeof(BProxJob<T>) * h.size(), hipMemcpyHostToDevice));
  B.stages.push_back(std::move(s));
}

// Room for every factor's compressed rows and columns: cap = max(k_b, n_b)
// entries (a sp projection keeps at most k_b; the default initial factors
// hold at most n_b).
template <class T>
static void sparse_alloc(Batch<T>& B, const std::vector<int64_t>& k) {
  const int nt = B.sp.nt, nf = B.sp.nf;
  auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
  std::vector<size_t> off((size_t)nt * nf);
  size_t tot = 0;
  for (int b = 0; b < nt; ++b) {
    const size_t n = B.sp.n[b], cap = (size_t)std::max<int64_t>(k[b], B.sp.n[b]);
    for (int f = 0; f < nf; ++f) {
      off[(size_t)b * nf + f] = tot;
      tot += 2 * up(sizeof(int) * (n + 1)) + 3 * up(sizeof(int) * cap) + 2 * up(sizeof(T) * cap) + up(sizeof(int));
    }
  }
  B.spmem.alloc(std::max<size_t>(tot, 1));
  // every byte 0xFF: the kept-entry counts start at -1 (no list yet)
  FH_HIP(hipMemset(B.spmem.p, 0xFF, std::max<size_t>(tot, 1)));
  char* base = static_cast<char*>(B.spmem.p);
  B.spfac_h.resize((size_t)nt * nf);
  for (int b = 0; b < nt; ++b) {
    const size_t n = B.sp.n[b], cap = (size_t)std::max<int64_t>(k[b], B.sp.n[b]);
    for (int f = 0; f < nf; ++f) {
      char* q = base + off[(size_t)b * nf + f];
      BSpFac<T> F;
      F.rp = reinterpret_cast<int*>(q), q += up(sizeof(int) * (n + 1));
      F.cp = reinterpret_cast<int*>(q), q += up(sizeof(int) * (n + 1));
      F.ci = reinterpret_cast<int*>(q), q += up(sizeof(int) * cap);
      F.ri = reinterpret_cast<int*>(q), q += up(sizeof(int) * cap);
      F.v = reinterpret_cast<T*>(q), q += up(sizeof(T) * cap);
      F.cv = reinterpret_cast<T*>(q), q += up(sizeof(T) * cap);
      F.klist = reinterpret_cast<int*>(q), q += up(sizeof(int) * cap);
      F.kcount = reinterpret_cast<int*>(q);
      F.cap = (int)cap;
      B.spfac_h[(size_t)b * nf + f] = F;
    }
  }
  B.spfac.alloc(sizeof(BSpFac<T>) * B.spfac_h.size());
  FH_HIP(hipMemcpy(B.spfac.p, B.spfac_h.data(), sizeof(BSpFac<T>) * B.spfac_h.size(), hipMemcpyHostToDevice));
}

template <class T>
static void fill_build_jobs(Batch<T>& B, const std::vector<int>& fs, typename Batch<T>::Stage& s) {
  std::vector<BSpBuildJob<T>> h;
  for (int b = 0; b < B.sp.nt; ++b)
    for (int f : fs) h.push_back(BSpBuildJob<T>{B.mat(b, B.sS(f)), B.spfac_h[(size_t)b * B.sp.nf + f], B.sp.n[b]});
  s.kind = 5;
  s.njobs = (int)h.size();
  s.jobs.alloc(sizeof(BSpBuildJob<T>) * h.size());
  FH_HIP(hipMemcpy(s.jobs.p, h.data(), sizeof(BSpBuildJob<T>) * h.size(), hipMemcpyHostToDevice));
}

template <class T>
static void add_sparse_build(Batch<T>& B, int f) {
  typename Batch<T>::Stage s;
  fill_build_jobs(B, {f}, s);
  B.stages.push_back(std::move(s));
}

// the compressed forms of every factor as they stand (before the first iteration)
template <class T>
static void sparse_build_all(Batch<T>& B, hipStream_t st) {
  typename Batch<T>::Stage s;
  std::vector<int> fs(B.sp.nf);
  std::iota(fs.begin(), fs.end(), 0);
  fill_build_jobs(B, fs, s);
  launch_bsparse<T>(s.jobs.template as<BSpBuildJob<T>>(), s.njobs, st);
  FH_HIP(hipStreamSynchronize(st));
}

// The R side (factors f+1 .. nf-1, norm to nR) and the L side (0 .. f-1, to
// nL) of factor f's update, in one launch: the two sides of every target are
// independent jobs (a side without factors writes 1 at once).
template <class T>
static void add_power_sp(Batch<T>& B, int f) {
  const int nf = B.sp.nf;
  std::vector<BPowSpJob<T>> h;
  const BSpFac<T>* fac = B.spfac.template as<BSpFac<T>>();
  for (int side = 0; side < 2; ++side)
    for (int b = 0; b < B.sp.nt; ++b) {
      const int first = side == 0 ? f + 1 : 0, m = side == 0 ? nf - 1 - f : f;
      double* out = (side == 0 ? B.nR : B.nL).template as<double>();
      h.push_back(BPowSpJob<T>{fac + (size_t)b * nf + first, m, B.sp.n[b], b, out,
                               B.powv.template as<double>() + B.powoff[b] + (int64_t)(2 * f + side) * B.sp.n[b]});
    }
  typename Batch<T>::Stage s;
  s.kind = 6;
  s.njobs = (int)h.size();
  int mc = -1;
  for (const auto& j : h) {
    s.nmax = std::max(s.nmax, j.n);
    if (j.m > 0) mc = (mc == -1 || mc == j.m) ? j.m : 0;
  }
  // FH_BATCH_POWER_REG=0 reads the rows from memory in every round instead
  const char* reg = std::getenv("FH_BATCH_POWER_REG");
  const bool reg_ok = reg == nullptr || std::string(reg) != "0";
  s.mc = (mc == 1 || mc == 2) && s.nmax <= 512 && reg_ok ? mc : 0;
  s.jobs.alloc(sizeof(BPowSpJob<T>) * h.size());
  FH_HIP(hipMemcpy(s.jobs.p, h.data(), sizeof(BPowSpJob<T>) * h.size(), hipMemcpyHostToDevice));
  B.stages.push_back(std::move(s));
}

template <class T>
static void add_lambda(Batch<T>& B) {
  std::vector<BLamJob<T>> h(B.sp.nt);
  for (int b = 0; b < B.sp.nt; ++b)
    h[b] = BLamJob<T>{B.mat(b, B.sA()), B.mat(b, B.sSM()), (int64_t)B.sp.n[b] * B.sp.n[b], b};
  typename Batch<T>::Stage s;
  s.kind = 4;
  s.njobs = B.sp.nt;
  s.jobs.alloc(sizeof(BLamJob<T>) * h.size());
  FH_HIP(hipMemcpy(s.jobs.p, h.data(), sizeof(BLamJob<T>) * h.size(), hipMemcpyHostToDevice));
  B.stages.push_back(std::move(s));
}

// The fixed schedule of one PALM iteration (L2R, packed partial products).
template <class T>
static void build_schedule(Batch<T>& B, const std::vector<int64_t>& k) {
  const int nf = B.sp.nf;
  B.stages.clear();
  auto nn = [&](int b) { return B.sp.n[b]; };
  auto job = [&](int b, const T* A, int ta, const T* Bm, int tb, const T* Cin, T* Cout, int am, int bm) {
    return BGemmJob<T>{A, Bm, Cin, Cout, nn(b), nn(b), nn(b), ta, tb, am, bm, b};
  };
  // packed right products of the start of the sweep: PR'[nf-2] = S[nf-1], PR[i] = S[i+1] PR'[i+1]
  auto PRp = [&](int b, int i) -> T* { return i == nf - 2 ? B.mat(b, B.sS(nf - 1)) : B.mat(b, B.sPR(i)); };
  auto PLp = [&](int b, int f) -> T* { return f == 1 ? B.mat(b, B.sS(0)) : B.mat(b, B.sPL(f)); };
  for (int i = nf - 3; i >= 0; --i)
    add_gemm<T>(B, [&, i](int b) { return job(b, B.mat(b, B.sS(i + 1)), 0, PRp(b, i + 1), 0, nullptr, B.mat(b, B.sPR(i)), 0, 0); });
  for (int f = 0; f < nf; ++f) {
    const bool hasR = f < nf - 1, hasL = f > 0;
    if (B.sparse) {
      add_power_sp<T>(B, f);
    } else {
      add_power<T>(B, 1, [&, f, hasR](int b) -> const T* { return hasR ? PRp(b, f) : nullptr; }, 2 * f);
      add_power<T>(B, 2, [&, f, hasL](int b) -> const T* { return hasL ? PLp(b, f) : nullptr; }, 2 * f + 1);
    }
    if (hasR && hasL) {
      add_gemm<T>(B, [&, f](int b) { return job(b, B.mat(b, B.sS(f)), 0, PRp(b, f), 0, nullptr, B.mat(b, B.sDR()), 0, 0); });
      add_gemm<T>(B, [&, f](int b) {
        return job(b, PLp(b, f), 0, B.mat(b, B.sDR()), 0, B.mat(b, B.sA()), B.mat(b, B.sTMP()), 1, 1);
      });
      add_gemm<T>(B, [&, f](int b) { return job(b, B.mat(b, B.sTMP()), 0, PRp(b, f), 1, nullptr, B.mat(b, B.sTR()), 0, 0); });
      add_gemm<T>(B, [&, f](int b) {
        return job(b, PLp(b, f), 1, B.mat(b, B.sTR()), 0, B.mat(b, B.sS(f)), B.mat(b, B.sS(f)), 2, 2);
      });
    } else if (hasR) {
      add_gemm<T>(B, [&, f](int b) {
        return job(b, B.mat(b, B.sS(f)), 0, PRp(b, f), 0, B.mat(b, B.sA()), B.mat(b, B.sTMP()), 1, 1);
      });
      add_gemm<T>(B, [&, f](int b) {
        return job(b, B.mat(b, B.sTMP()), 0, PRp(b, f), 1, B.mat(b, B.sS(f)), B.mat(b, B.sS(f)), 2, 2);
      });
    } else {
      add_gemm<T>(B, [&, f](int b) {
        return job(b, PLp(b, f), 0, B.mat(b, B.sS(f)), 0, B.mat(b, B.sA()), B.mat(b, B.sTMP()), 1, 1);
      });
      add_gemm<T>(B, [&, f](int b) {
        return job(b, PLp(b, f), 1, B.mat(b, B.sTMP()), 0, B.mat(b, B.sS(f)), B.mat(b, B.sS(f)), 2, 2);
      });
    }
    add_prox<T>(B, f, k);
    if (B.sparse) add_sparse_build<T>(B, f);
    if (f >= 1 && f < nf - 1)
      add_gemm<T>(B, [&, f](int b) { return job(b, PLp(b, f), 0, B.mat(b, B.sS(f)), 0, nullptr, B.mat(b, B.sPL(f + 1)), 0, 0); });
  }
  // update_lambda with the packed left product of the sweep: S = PL'[nf-1] S[nf-1]
  add_gemm<T>(B, [&](int b) { return job(b, PLp(b, nf - 1), 0, B.mat(b, B.sS(nf - 1)), 0, nullptr, B.mat(b, B.sSM()), 0, 0); });
  add_lambda<T>(B);
}

template <class T>
static void run_stage(Batch<T>& B, const typename Batch<T>::Stage& s, hipStream_t st) {
  BScal sc{B.lam.template as<double>(), B.nR.template as<double>(), B.nL.template as<double>()};
  switch (s.kind) {
    case 0:
      launch_bgemm<T>(s.jobs.template as<BGemmJob<T>>(), B.tiles32.template as<int3>(), B.nt32, 32, sc, st);
      launch_bgemm<T>(s.jobs.template as<BGemmJob<T>>(), B.tiles64.template as<int3>(), B.nt64, 64, sc, st);
      launch_bgemm<T>(s.jobs.template as<BGemmJob<T>>(), B.tilesM.template as<int3>(), B.ntM, BGEMM_MFMA, sc, st);
      break;
    case 1:
    case 2:
      launch_bpower<T>(s.jobs.template as<BPowJob<T>>(), s.njobs, B.maxit, B.thr, B.warm, st);
      break;
    case 3:
      launch_bprox<T>(s.jobs.template as<BProxJob<T>>(), s.njobs, B.prox_mode, st);
      break;
    case 4:
      launch_blambda<T>(s.jobs.template as<BLamJob<T>>(), s.njobs, B.lam.template as<double>(), st);
      break;
    case 5:
      launch_bsparse<T>(s.jobs.template as<BSpBuildJob<T>>(), s.njobs, st);
      break;
    case 6:
    case 7:
      launch_bpower_sp<T>(s.jobs.template as<BPowSpJob<T>>(), s.njobs, s.nmax, s.mc, B.maxit, B.thr, B.warm, st);
      break;
    case 8:
      launch_bspmm<T>(s.jobs.template as<BSpmmJob<T>>(), s.tiles.template as<int3>(), s.ntiles, sc, st);
      break;
  }
}

template <class T>
static void run_iteration(Batch<T>& B, hipStream_t st) {
  for (auto& s : B.stages) run_stage(B, s, st);
}

// Autotuner: for every distinct target size, time one full-size GEMM stage
// with 32- and 64-wide VALU tiles and (FP32) the 64 x 64 MFMA tile, and keep
// the fastest.
template <class T>
static void autotune(Batch<T>& B, hipStream_t st) {
  std::map<int, std::vector<int>> by_n;
  for (int b = 0; b < B.sp.nt; ++b) by_n[B.sp.n[b]].push_back(b);
  B.tile_of.assign(B.sp.nt, 64);
  // FH_BGEMM_TILE=32 / 64 / 65 (MFMA, FP32 only) fixes the choice
  if (const char* e = getenv("FH_BGEMM_TILE")) {
    const int code = atoi(e);
    if (code == 32 || code == 64 || (code == BGEMM_MFMA && std::is_same<T, float>::value)) {
      B.tile_of.assign(B.sp.nt, code);
      return;
    }
  }
  // a GEMM stage over the members of one size class
  const typename Batch<T>::Stage* g = nullptr;
  for (auto& s : B.stages)
    if (s.kind == 0) { g = &s; break; }
  if (!g) return;
  hipEvent_t e0, e1;
  FH_HIP(hipEventCreate(&e0));
  FH_HIP(hipEventCreate(&e1));
  BScal sc{B.lam.template as<double>(), B.nR.template as<double>(), B.nL.template as<double>()};
  for (auto& kv : by_n) {
    float best = 1e30f;
    int best_tile = 64;
    std::vector<int> cands{32, 64};
    if (std::is_same<T, float>::value) cands.push_back(BGEMM_MFMA);
    for (int code : cands) {
      const int tile = code == 32 ? 32 : 64;
      std::vector<int3> tl;
      for (int b : kv.second) {
        const int nt = (kv.first + tile - 1) / tile;
        for (int i = 0; i < nt; ++i)
          for (int j = 0; j < nt; ++j) tl.push_back(make_int3(b, i, j));
      }
      DBuf dt(sizeof(int3) * tl.size());
      FH_HIP(hipMemcpy(dt.p, tl.data(), sizeof(int3) * tl.size(), hipMemcpyHostToDevice));
      // the fastest of 3 timed launches, so that one slow launch does not decide the tile
      float ms = 1e30f;
      for (int rep = 0; rep < 3; ++rep) {
        FH_HIP(hipEventRecord(e0, st));
        launch_bgemm<T>(g->jobs.template as<BGemmJob<T>>(), dt.as<int3>(), (int)tl.size(), code, sc, st);
        FH_HIP(hipEventRecord(e1, st));
        FH_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        FH_HIP(hipEventElapsedTime(&t, e0, e1));
        ms = std::min(ms, t);
      }
      if (ms < best) {
        best = ms;
        best_tile = code;
      }
    }
    for (int b : kv.second) B.tile_of[b] = best_tile;
  }
  FH_HIP(hipEventDestroy(e0));
  FH_HIP(hipEventDestroy(e1));
}

template <class T>
static double run_its(Batch<T>& B, int its, bool graph, hipStream_t st) {
  auto t0 = std::chrono::steady_clock::now();
  if (its <= 0) return 0.0;
  if (graph) {
    hipGraph_t g;
    hipGraphExec_t ge;
    FH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      run_iteration(B, st);
    } catch (...) {
      // never leave the stream capturing: end (and drop) the partial graph first
      if (hipStreamEndCapture(st, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
      throw;
    }
    FH_HIP(hipStreamEndCapture(st, &g));
    FH_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < its; ++i) FH_HIP(hipGraphLaunch(ge, st));
    FH_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    FH_HIP(hipGraphExecDestroy(ge));
    FH_HIP(hipGraphDestroy(g));
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
  for (int i = 0; i < its; ++i) run_iteration(B, st);
  FH_HIP(hipStreamSynchronize(st));
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// FP64 relative error of lambda * S_0 ... S_{nf-1} against A, per target.
// The scratch FP64 batch and job lists are built once and reused by every check.
struct Verifier {
  BatchSpec sp;
  Batch<double> V;
  std::vector<DBuf> gemm_jobs;  // one job list per factor index
  DBuf lam_jobs, out;
  bool ready = false;

  void init(const BatchSpec& s, const std::vector<const double*>& Ad, hipStream_t st) {
    sp = s;
    BatchSpec vs = s;
    vs.nf = 1;
    batch_alloc(V, vs, false);
    V.tile_of.assign(s.nt, 64);
    build_tiles(V);
    // V slots (nf = 1): 0 A, 1 running product, 2 factor cast to FP64, 3 GEMM output
    for (int b = 0; b < s.nt; ++b) {
      const int64_t len = (int64_t)s.n[b] * s.n[b];
      FH_HIP(hipMemcpyAsync(V.mat(b, 0), Ad[b], sizeof(double) * len, hipMemcpyDeviceToDevice, st));
    }
    std::vector<BGemmJob<double>> h(s.nt);
    for (int b = 0; b < s.nt; ++b)
      h[b] = BGemmJob<double>{V.mat(b, 2), V.mat(b, 1), nullptr, V.mat(b, 3), s.n[b], s.n[b], s.n[b], 0, 0, 0, 0, b};
    gemm_jobs.resize(1);
    gemm_jobs[0].alloc(sizeof(BGemmJob<double>) * s.nt);
    FH_HIP(hipMemcpy(gemm_jobs[0].p, h.data(), sizeof(BGemmJob<double>) * s.nt, hipMemcpyHostToDevice));
    std::vector<BLamJob<double>> hl(s.nt);
    for (int b = 0; b < s.nt; ++b) hl[b] = BLamJob<double>{V.mat(b, 0), V.mat(b, 1), (int64_t)s.n[b] * s.n[b], b};
    lam_jobs.alloc(sizeof(BLamJob<double>) * s.nt);
    FH_HIP(hipMemcpy(lam_jobs.p, hl.data(), sizeof(BLamJob<double>) * s.nt, hipMemcpyHostToDevice));
    out.alloc(sizeof(double) * s.nt);
    ready = true;
  }

  // per batch verified (its address): the cast job lists of every factor and
  // the copy of the GEMM output back to the running product, one launch each
  const void* cast_for = nullptr;
  std::vector<DBuf> cast_jobs;  // [f]: factor f of every target -> slot 2 (slot 1 for the last factor)
  DBuf copy_jobs;

  template <class T>
  void build_casts(Batch<T>& B) {
    cast_jobs.clear();
    cast_jobs.resize(sp.nf);
    std::vector<BCastJob> h(sp.nt);
    for (int f = 0; f < sp.nf; ++f) {
      for (int b = 0; b < sp.nt; ++b)
        h[b] = BCastJob{V.mat(b, f == sp.nf - 1 ? 1 : 2), B.mat(b, B.sS(f)), (int64_t)sp.n[b] * sp.n[b]};
      cast_jobs[f].alloc(sizeof(BCastJob) * sp.nt);
      FH_HIP(hipMemcpy(cast_jobs[f].p, h.data(), sizeof(BCastJob) * sp.nt, hipMemcpyHostToDevice));
    }
    for (int b = 0; b < sp.nt; ++b) h[b] = BCastJob{V.mat(b, 1), V.mat(b, 3), (int64_t)sp.n[b] * sp.n[b]};
    copy_jobs.alloc(sizeof(BCastJob) * sp.nt);
    FH_HIP(hipMemcpy(copy_jobs.p, h.data(), sizeof(BCastJob) * sp.nt, hipMemcpyHostToDevice));
    cast_for = &B;
  }

  template <class T>
  void run(Batch<T>& B, double* host_out, hipStream_t st) {
    if (cast_for != &B) build_casts(B);
    constexpr bool f32 = std::is_same<T, float>::value;
    BScal sc{V.lam.as<double>(), V.nR.as<double>(), V.nL.as<double>()};
    launch_bcast(cast_jobs[sp.nf - 1].as<BCastJob>(), sp.nt, f32, st);
    for (int f = sp.nf - 2; f >= 0; --f) {
      launch_bcast(cast_jobs[f].as<BCastJob>(), sp.nt, f32, st);
      launch_bgemm<double>(gemm_jobs[0].as<BGemmJob<double>>(), V.tiles64.as<int3>(), V.nt64, 64, sc, st);
      launch_bcast(copy_jobs.as<BCastJob>(), sp.nt, false, st);
    }
    launch_brelerr<double>(lam_jobs.as<BLamJob<double>>(), sp.nt, B.lam.template as<double>(), out.as<double>(), st);
    FH_HIP(hipMemcpyAsync(host_out, out.p, sizeof(double) * sp.nt, hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
  }
};

// Runs `its` iterations, in chunks of check_every with an FP64 check after
// each chunk when check_every > 0.  first_hit[b] records the iteration count
// (it0 + iterations run) at which target b first reached target_err.
// Returns true when every target has reached it.
struct Progress {
  std::vector<int> first_hit;
  int done = 0;
  double ms_all = -1.0;
};

template <class T>
static bool run_checked(Batch<T>& B, int its, int it0, bool graph, int check_every, const std::vector<double>& thr,
                        Verifier& ver, Progress& pr, std::chrono::steady_clock::time_point t_start,
                        double& ms_iter, hipStream_t st) {
  if (check_every <= 0) {
    ms_iter += run_its(B, its, graph, st);
    return false;
  }
  std::vector<double> err(B.sp.nt);
  int ran = 0;
  while (ran < its) {
    const int chunk = std::min(check_every, its - ran);
    ms_iter += run_its(B, chunk, graph, st);
    ran += chunk;
    ver.run(B, err.data(), st);
    for (int b = 0; b < B.sp.nt; ++b)
      if (pr.first_hit[b] < 0 && err[b] <= thr[b]) {
        pr.first_hit[b] = it0 + ran;
        pr.done++;
      }
    if (pr.done == B.sp.nt) {
      pr.ms_all = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
      return true;
    }
  }
  return false;
}

template <class T>
static void verify_fp64(const BatchSpec& sp, const std::vector<const double*>& Ad, Batch<T>& B, double* out,
                        hipStream_t st) {
  Verifier v;
  v.init(sp, Ad, st);
  v.run(B, out, st);
}

}  // namespace fh

struct fh_batch_result_s {
  int nt = 0, nf = 0;
  std::vector<int> n;
  std::vector<double> lambda, relerr;
  std::vector<std::vector<double>> fac;  // [target*nf + f] dense col-major
  std::vector<int> iters_to_target;      // -1: target error not reached (or not checked)
  double ms[5] = {0, 0, 0, 0, -1};       // setup, main iterations, refinement, verification, time to all targets
};

using namespace fh;

namespace {

template <class T>
void load_batch(Batch<T>& B, const std::vector<const double*>& Ad, const double* const* init,
                const double* lambda_init, hipStream_t st) {
  const int nf = B.sp.nf;
  std::vector<double> lam(B.sp.nt, 1.0);
  for (int b = 0; b < B.sp.nt; ++b) {
    const int n = B.sp.n[b];
    const int64_t len = (int64_t)n * n;
    // A (already FP64 on the device in Ad)
    if constexpr (std::is_same<T, float>::value) {
      std::vector<double> h(len);
      FH_HIP(hipMemcpy(h.data(), Ad[b], sizeof(double) * len, hipMemcpyDeviceToHost));
      std::vector<float> hf(h.begin(), h.end());
      FH_HIP(hipMemcpy(B.mat(b, B.sA()), hf.data(), sizeof(float) * len, hipMemcpyHostToDevice));
    } else {
      FH_HIP(hipMemcpy(B.mat(b, B.sA()), Ad[b], sizeof(double) * len, hipMemcpyDeviceToDevice));
    }
    for (int f = 0; f < nf; ++f) {
      std::vector<T> h(len, T(0));
      const double* src = init ? init[(size_t)b * nf + f] : nullptr;
      if (src) {
        for (int64_t i = 0; i < len; ++i) h[i] = (T)src[i];
      } else if (f > 0) {
        // palm4msa2 default init, left-to-right: S_0 = 0, the others identity
        for (int i = 0; i < n; ++i) h[(int64_t)i * n + i] = T(1);
      }
      FH_HIP(hipMemcpy(B.mat(b, B.sS(f)), h.data(), sizeof(T) * len, hipMemcpyHostToDevice));
    }
    if (lambda_init) lam[b] = lambda_init[b];
  }
  FH_HIP(hipMemcpy(B.lam.p, lam.data(), sizeof(double) * B.sp.nt, hipMemcpyHostToDevice));
}

}  // namespace

extern "C" int fh_palm4msa_batched(int ntargets, int nfacts, const int* n, const double* const* A, const int64_t* k,
                                   const double* const* supp, const double* const* init, const double* lambda_init,
                                   const fh_batch_params* p, fh_batch_result_t* out) {
  FH_API_BEGIN
  FH_CHECK(ntargets > 0 && nfacts >= 2 && n && A && p && out, "bad argument");
  FH_CHECK(p->dtype == FH_F32 || p->dtype == FH_F64, "batched PALM runs in FP32 or FP64");
  FH_CHECK(p->prox == FH_PROX_SP || p->prox == FH_PROX_SUPP, "batched PALM: prox must be sp or supp");
  FH_CHECK(p->prox == FH_PROX_SUPP ? supp != nullptr : k != nullptr, "batched PALM: missing k or supports");
  hipStream_t st;
  FH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  // pool blocks freed by earlier calls are reused in the order of the caller's
  // stream; this call runs on its own stream, so let that earlier work finish
  FH_HIP(hipDeviceSynchronize());
  auto r = std::make_unique<fh_batch_result_s>();
  auto t0 = std::chrono::steady_clock::now();
  BatchSpec sp{ntargets, nfacts, std::vector<int>(n, n + ntargets)};
  // targets in FP64 on the device (for the refinement and the verification)
  std::vector<DBuf> Ad(ntargets);
  std::vector<const double*> Adp(ntargets);
  for (int b = 0; b < ntargets; ++b) {
    const int64_t len = (int64_t)n[b] * n[b];
    Ad[b].alloc(sizeof(double) * len);
    FH_HIP(hipMemcpy(Ad[b].p, A[b], sizeof(double) * len, hipMemcpyHostToDevice));
    Adp[b] = Ad[b].as<double>();
  }
  std::vector<int64_t> kv;
  if (p->prox == FH_PROX_SP) kv.assign(k, k + ntargets);
  std::vector<double> thr(ntargets, p->target_err);  // error each target has to reach
  if (p->target_errs) thr.assign(p->target_errs, p->target_errs + ntargets);
  std::vector<double> lam_out(ntargets);
  if (p->dtype == FH_F32) {
    Batch<float> B;
    batch_alloc(B, sp, p->prox == FH_PROX_SUPP);
    B.prox_mode = p->prox == FH_PROX_SUPP ? 1 : 0;
    B.maxit = p->norm2_max_iter > 0 ? p->norm2_max_iter : 100;
    B.thr = p->norm2_threshold > 0 ? p->norm2_threshold : 1e-6;
    B.warm = p->warm_power;
    load_batch(B, Adp, init, lambda_init, st);
    if (B.prox_mode == 1)
      for (int b = 0; b < ntargets; ++b)
        for (int f = 0; f < nfacts; ++f) {
          const int64_t len = (int64_t)n[b] * n[b];
          std::vector<float> h(supp[(size_t)b * nfacts + f], supp[(size_t)b * nfacts + f] + len);
          FH_HIP(hipMemcpy(B.mat(b, B.sMASK(f)), h.data(), sizeof(float) * len, hipMemcpyHostToDevice));
        }
    // sp projections keeping at most 1/8 of each factor: step sizes from the
    // sparse factors (FH_BATCH_SPARSE_POWER=0 keeps the dense products)
    {
      const char* e = getenv("FH_BATCH_SPARSE_POWER");
      bool sparse = B.prox_mode == 0 && !(e && e[0] == '0');
      for (int b = 0; b < ntargets && sparse; ++b) sparse = kv[b] <= (int64_t)n[b] * n[b] / 8;
      B.sparse = sparse;
      if (sparse) sparse_alloc(B, kv);
    }
    B.tile_of.assign(ntargets, 64);
    build_tiles(B);
    build_schedule(B, kv);
    if (B.sparse) sparse_build_all(B, st);
    if (p->autotune) {
      autotune(B, st);
      build_tiles(B);
    }
    r->ms[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    Verifier ver;
    Progress pr;
    pr.first_hit.assign(ntargets, -1);
    if (p->check_every > 0) ver.init(sp, Adp, st);
    const auto t_start = std::chrono::steady_clock::now();
    const bool all = run_checked(B, p->num_its, 0, p->use_graph != 0, p->check_every, thr, ver, pr,
                                 t_start, r->ms[1], st);
    // frozen-support FP64 refinement
    Batch<double> D;
    if (p->refine_its > 0 && !all) {
      auto t1 = std::chrono::steady_clock::now();
      batch_alloc(D, sp, true);
      D.prox_mode = 1;
      D.maxit = B.maxit;
      D.thr = B.thr;
      D.warm = p->warm_power;
      for (int b = 0; b < ntargets; ++b) {
        const int64_t len = (int64_t)n[b] * n[b];
        FH_HIP(hipMemcpyAsync(D.mat(b, D.sA()), Adp[b], sizeof(double) * len, hipMemcpyDeviceToDevice, st));
        for (int f = 0; f < nfacts; ++f) {
          launch_cast_f32_f64(D.mat(b, D.sS(f)), B.mat(b, B.sS(f)), len, st);
          launch_mask_nonzero_f32(D.mat(b, D.sMASK(f)), B.mat(b, B.sS(f)), len, st);
        }
      }
      FH_HIP(hipMemcpyAsync(D.lam.p, B.lam.p, sizeof(double) * ntargets, hipMemcpyDeviceToDevice, st));
      FH_HIP(hipMemcpyAsync(D.powv.p, B.powv.p, B.powv.bytes, hipMemcpyDeviceToDevice, st));
      D.tile_of = B.tile_of;  // the FP32 MFMA tile has no FP64 form: its targets take 64-wide VALU tiles
      for (int& t : D.tile_of)
        if (t == BGEMM_MFMA) t = 64;
      build_tiles(D);
      build_schedule(D, {});
      FH_HIP(hipStreamSynchronize(st));
      const double setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
      run_checked(D, p->refine_its, p->num_its, p->use_graph != 0, p->check_every, thr, ver, pr, t_start,
                  r->ms[2], st);
      r->ms[0] += setup_ms;
      auto t2 = std::chrono::steady_clock::now();
      r->relerr.resize(ntargets);
      verify_fp64(sp, Adp, D, r->relerr.data(), st);
      r->ms[3] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t2).count();
      FH_HIP(hipMemcpy(lam_out.data(), D.lam.p, sizeof(double) * ntargets, hipMemcpyDeviceToHost));
      r->fac.resize((size_t)ntargets * nfacts);
      for (int b = 0; b < ntargets; ++b)
        for (int f = 0; f < nfacts; ++f) {
          const int64_t len = (int64_t)n[b] * n[b];
          r->fac[(size_t)b * nfacts + f].resize(len);
          FH_HIP(hipMemcpy(r->fac[(size_t)b * nfacts + f].data(), D.mat(b, D.sS(f)), sizeof(double) * len,
                           hipMemcpyDeviceToHost));
        }
    } else {
      auto t2 = std::chrono::steady_clock::now();
      r->relerr.resize(ntargets);
      verify_fp64(sp, Adp, B, r->relerr.data(), st);
      r->ms[3] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t2).count();
      FH_HIP(hipMemcpy(lam_out.data(), B.lam.p, sizeof(double) * ntargets, hipMemcpyDeviceToHost));
      r->fac.resize((size_t)ntargets * nfacts);
      for (int b = 0; b < ntargets; ++b)
        for (int f = 0; f < nfacts; ++f) {
          const int64_t len = (int64_t)n[b] * n[b];
          std::vector<float> h(len);
          FH_HIP(hipMemcpy(h.data(), B.mat(b, B.sS(f)), sizeof(float) * len, hipMemcpyDeviceToHost));
          r->fac[(size_t)b * nfacts + f].assign(h.begin(), h.end());
        }
    }
    r->iters_to_target = pr.first_hit;
    r->ms[4] = pr.ms_all;
  } else {
    Batch<double> B;
    batch_alloc(B, sp, p->prox == FH_PROX_SUPP);
    B.prox_mode = p->prox == FH_PROX_SUPP ? 1 : 0;
    B.maxit = p->norm2_max_iter > 0 ? p->norm2_max_iter : 100;
    B.thr = p->norm2_threshold > 0 ? p->norm2_threshold : 1e-6;
    B.warm = p->warm_power;
    load_batch(B, Adp, init, lambda_init, st);
    if (B.prox_mode == 1)
      for (int b = 0; b < ntargets; ++b)
        for (int f = 0; f < nfacts; ++f)
          FH_HIP(hipMemcpy(B.mat(b, B.sMASK(f)), supp[(size_t)b * nfacts + f], sizeof(double) * n[b] * n[b],
                           hipMemcpyHostToDevice));
    B.tile_of.assign(ntargets, 64);
    build_tiles(B);
    build_schedule(B, kv);
    if (p->autotune) {
      autotune(B, st);
      build_tiles(B);
    }
    r->ms[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    Verifier ver;
    Progress pr;
    pr.first_hit.assign(ntargets, -1);
    if (p->check_every > 0) ver.init(sp, Adp, st);
    const auto t_start = std::chrono::steady_clock::now();
    run_checked(B, p->num_its + p->refine_its, 0, p->use_graph != 0, p->check_every, thr, ver, pr, t_start,
                r->ms[1], st);
    r->iters_to_target = pr.first_hit;
    r->ms[4] = pr.ms_all;
    auto t2 = std::chrono::steady_clock::now();
    r->relerr.resize(ntargets);
    verify_fp64(sp, Adp, B, r->relerr.data(), st);
    r->ms[3] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t2).count();
    FH_HIP(hipMemcpy(lam_out.data(), B.lam.p, sizeof(double) * ntargets, hipMemcpyDeviceToHost));
    r->fac.resize((size_t)ntargets * nfacts);
    for (int b = 0; b < ntargets; ++b)
      for (int f = 0; f < nfacts; ++f) {
        const int64_t len = (int64_t)n[b] * n[b];
        r->fac[(size_t)b * nfacts + f].resize(len);
        FH_HIP(hipMemcpy(r->fac[(size_t)b * nfacts + f].data(), B.mat(b, B.sS(f)), sizeof(double) * len,
                         hipMemcpyDeviceToHost));
      }
  }
  if (r->iters_to_target.empty()) r->iters_to_target.assign(ntargets, -1);
  r->nt = ntargets;
  r->nf = nfacts;
  r->n = sp.n;
  r->lambda = lam_out;
  FH_HIP(hipStreamDestroy(st));
  *out = r.release();
  FH_API_END
}

extern "C" int fh_batch_result_lambda(fh_batch_result_t r, double* out) {
  FH_API_BEGIN
  FH_CHECK(r && out, "null argument");
  std::copy(r->lambda.begin(), r->lambda.end(), out);
  FH_API_END
}

extern "C" int fh_batch_result_relerr(fh_batch_result_t r, double* out) {
  FH_API_BEGIN
  FH_CHECK(r && out, "null argument");
  std::copy(r->relerr.begin(), r->relerr.end(), out);
  FH_API_END
}

extern "C" int fh_batch_result_factor(fh_batch_result_t r, int target, int f, double* out) {
  FH_API_BEGIN
  FH_CHECK(r && out && target >= 0 && target < r->nt && f >= 0 && f < r->nf, "bad index");
  const auto& v = r->fac[(size_t)target * r->nf + f];
  std::copy(v.begin(), v.end(), out);
  FH_API_END
}

extern "C" int fh_batch_result_times(fh_batch_result_t r, double* ms5) {
  FH_API_BEGIN
  FH_CHECK(r && ms5, "null argument");
  for (int i = 0; i < 5; ++i) ms5[i] = r->ms[i];
  FH_API_END
}

extern "C" int fh_batch_result_iters(fh_batch_result_t r, int* iters) {
  FH_API_BEGIN
  FH_CHECK(r && iters, "null argument");
  std::copy(r->iters_to_target.begin(), r->iters_to_target.end(), iters);
  FH_API_END
}

extern "C" int fh_batch_result_free(fh_batch_result_t r) {
  FH_API_BEGIN
  delete r;
  FH_API_END
}
