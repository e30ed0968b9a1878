// fh_palm.cpp -- PALM4MSA 2020 and hierarchical drivers on the device (FP64).
//
// Host C++ orchestration of the device kernels, following
//   palm4msa2        src/algorithm/factorization/faust_palm4msa2020.hpp:47-300
//   update_fact      faust_palm4msa2020.hpp:528-676
//   grad2            faust_palm4msa2020.hpp:363-439 (the CPU path: dense D)
//   grad1            faust_palm4msa2020.hpp:304-361 (USE_GRAD1 / the GPU2 path)
//   update_lambda    faust_palm4msa2020.hpp:442-493
//   hierarchical     src/algorithm/factorization/faust_hierarchical.hpp:1-150
// Factors are shared_ptr device matrices so that the left/right partial
// products keep the factor objects they were built from, as the reference's
// TransformHelper pointer vectors do (update_lambda then sees the left
// products of the start of the sweep).  All products go through matmul_f64,
// whose results are independent of the operand formats.
#include <algorithm>
#include <climits>
#include <functional>

#include "fh_api_common.hpp"

namespace fh {

using Chain = std::vector<DMatP>;

struct TrajPoint {
  double call, it, relerr, lambda;
};

static DMatP dmat_zero_sparse(int m, int n, hipStream_t st) {
  std::vector<int32_t> rp(m + 1, 0);
  return make_csr_from_host(FH_F64, m, n, 0, rp.data(), nullptr, nullptr, st);
}
static DMatP dmat_eye(int m, int n, bool sparse, hipStream_t st) {
  int mn = std::min(m, n);
  if (sparse) {
    std::vector<int32_t> rp(m + 1, 0), ci(mn);
    std::vector<double> v(mn, 1.0);
    for (int i = 0; i < m; ++i) rp[i + 1] = rp[i] + (i < mn ? 1 : 0);
    for (int i = 0; i < mn; ++i) ci[i] = i;
    auto M = make_csr_from_host(FH_F64, m, n, mn, rp.data(), ci.data(), v.data(), st);
    M->is_identity = true;
    return M;
  }
  auto M = make_dense(FH_F64, m, n, nullptr, std::max(m, 1), false, st);
  eye_f64(M->d.as<double>(), m, n, st);
  M->nnz = mn;
  M->is_identity = true;
  return M;
}
static DMatP dmat_zero_dense(int m, int n, hipStream_t st) {
  auto M = make_dense(FH_F64, m, n, nullptr, std::max(m, 1), false, st);
  fill_f64(M->d.as<double>(), (int64_t)m * n, 0.0, st);
  M->nnz = 0;
  return M;
}

static DMatP to_sparse(const DMat& D, hipStream_t st) {
  FH_CHECK(D.kind == K_DENSE, "to_sparse: dense input expected");
  auto S = std::make_shared<DMat>();
  dense_to_csr_f64(D.d.as<double>(), D.m, D.n, *S, st);
  return S;
}

// ------------------------------------------------------------ dynprog packing
// faust_prod_opt.cpp:6-86 costs in wrapping int32 arithmetic
static int32_t imul(int64_t a, int64_t b) { return (int32_t)(uint32_t)((uint64_t)a * (uint64_t)b); }
static int32_t iadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

// dynprog_multiply (faust_prod_opt_gen.hpp:81-200), op 'N'
static DMatP dynprog_multiply(const Chain& f, hipStream_t st) {
  const int n = (int)f.size();
  if (n == 1) return dense_of(*f[0], st);
  std::vector<std::vector<int32_t>> c(n, std::vector<int32_t>(n, 0));
  std::vector<std::vector<int>> inds(n, std::vector<int>(n, 0));
  auto nnz = [&](int i) { return (int)nnz_of(*f[i], st); };
  for (int p = 1; p < n; ++p)
    for (int i = 0; i < n - p; ++i) {
      const int j = p + i;
      const bool ld = f[i]->kind == K_DENSE, rd = f[j]->kind == K_DENSE;
      c[i][j] = INT_MAX;
      for (int k = i; k < j; ++k) {
        int32_t cost = iadd(c[i][k], c[k + 1][j]);
        const int mf_nc = f[k]->n;
        int32_t add;
        if (ld && rd)
          add = imul(imul(f[i]->m, mf_nc), f[j]->n);
        else if (ld)
          add = imul(f[i]->m, nnz(j));
        else if (rd)
          add = imul(nnz(i), f[j]->n);
        else
          add = imul(nnz(i), nnz(j)) / f[i]->m;
        cost = iadd(cost, add);
        if (cost < c[i][j]) {
          c[i][j] = cost;
          inds[i][j] = k;
        }
      }
    }
  std::function<DMatP(int, int)> rec = [&](int i, int j) -> DMatP {
    if (i == j) return f[i];
    auto out = std::make_shared<DMat>();
    if (j - i == 1) {
      matmul_f64(*f[i], 'N', *f[j], 'N', 1.0, 0.0, *out, st);
      return out;
    }
    const int k = inds[i][j];
    DMatP l = rec(i, k), r = rec(k + 1, j);
    matmul_f64(*l, 'N', *r, 'N', 1.0, 0.0, *out, st);
    return out;
  };
  return rec(0, n - 1);
}

// pack_factors(DYNPROG) of a whole chain (faust_TransformHelper.hpp:1666-1716)
static Chain packed(const Chain& c, hipStream_t st) {
  if (c.empty()) return c;
  if (c.size() == 1) {
    if (c[0]->kind == K_DENSE) return c;
    return Chain{dense_of(*c[0], st)};
  }
  return Chain{dynprog_multiply(c, st)};
}

// ------------------------------------------------------------ PALM pieces
static double read_scalar(const DBuf& b, hipStream_t st) {
  double h;
  FH_HIP(hipMemcpyAsync(&h, b.p, sizeof(double), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  return h;
}

static double fro_norm(const DMat& M, hipStream_t st) {
  DMatP D;
  const DMat* p = &M;
  if (M.kind != K_DENSE) {
    D = dense_of(M, st);
    p = D.get();
  }
  DBuf s(sizeof(double));
  reduce_dot_f64(p->d.as<double>(), p->d.as<double>(), (int64_t)p->m * p->n, s.as<double>(), st);
  return std::sqrt(read_scalar(s, st));
}

// multiply_order_opt_all_ends, MatGeneric version (faust_prod_opt.hpp:49-86)
static void mul_all_ends(std::vector<const DMat*> facts, DMat& out, double alpha, double beta, std::vector<char> fl,
                         hipStream_t st) {
  const int nf = (int)facts.size();
  int ri = 0, li = nf - 1;
  DMat tmpr, tmpl;
  auto flag = [&](int i) { return fl[(int)fl.size() > i ? i : 0]; };
  while (li - ri > 1) {
    const DMat *R1 = facts[ri], *R2 = facts[ri + 1], *L1 = facts[li - 1], *L2 = facts[li];
    const int64_t lhs = (int64_t)R1->m * R1->n * R2->n, rhs = (int64_t)L1->m * L1->n * L2->n;
    if (lhs < rhs) {
      DMat t;
      matmul_f64(*R1, flag(ri), *R2, flag(ri + 1), 1.0, 0.0, t, st);
      tmpr = std::move(t);
      ri++;
      facts[ri] = &tmpr;
      if ((int)fl.size() > ri) fl[ri] = 'N';
    } else {
      DMat t;
      matmul_f64(*L1, flag(li - 1), *L2, flag(li), 1.0, 0.0, t, st);
      tmpl = std::move(t);
      li--;
      facts[li] = &tmpl;
      if ((int)fl.size() > li) fl[li] = 'N';
    }
  }
  const char opl = ri == 0 ? fl[0] : 'N';
  const char opr = (li == nf - 1 && (int)fl.size() > li) ? fl[li] : 'N';
  matmul_f64(*facts[ri], opl, *facts[li], opr, alpha, beta, out, st);
}

static DMatP copy_dense(const DMat& A, hipStream_t st) { return dense_of(A, st); }

// compute_n_apply_grad2: D <- D - (lambda/c) L^T (lambda L D R - A) R^T
// Dp is the current factor in its stored format (used for the products, whose
// bits do not depend on the format); D is its dense copy, updated in place.
static void grad2(const Chain& pL, const Chain& pR, const DMat& A, double lambda, double c, const DMat& Dp, DMat& D,
                  const fh_stop_crit& sc, double& error, hipStream_t st) {
  DMatP tmp = copy_dense(A, st);
  const DMat *L = nullptr, *R = nullptr;
  DMatP Ld, Rd;
  if (!pR.empty()) {
    if (pR.size() == 1)
      R = pR[0].get();
    else {
      Rd = faust_get_product_f64(pR, st);
      R = Rd.get();
    }
  }
  if (!pL.empty()) {
    if (pL.size() == 1)
      L = pL[0].get();
    else {
      Ld = faust_get_product_f64(pL, st);
      L = Ld.get();
    }
  }
  std::vector<const DMat*> facts;
  std::vector<char> fl;
  if (R && L) {
    mul_all_ends({L, &Dp, R}, *tmp, lambda, -1.0, {'N'}, st);
    if (sc.is_criterion_error) error = fro_norm(*tmp, st);
    facts = {L, tmp.get(), R};
    fl = {'T', 'N', 'T'};
  } else if (R) {
    mul_all_ends({&Dp, R}, *tmp, lambda, -1.0, {'N'}, st);
    if (sc.is_criterion_error) error = fro_norm(*tmp, st);
    facts = {tmp.get(), R};
    fl = {'N', 'T'};
  } else {
    mul_all_ends({L, &Dp}, *tmp, lambda, -1.0, {'N'}, st);
    if (sc.is_criterion_error) error = fro_norm(*tmp, st);
    facts = {L, tmp.get()};
    fl = {'T', 'N'};
  }
  DMat g;
  mul_all_ends(facts, g, lambda / c, 0.0, fl, st);
  sub_f64(D.d.as<double>(), g.d.as<double>(), (int64_t)D.m * D.n, st);
}

// compute_n_apply_grad1 (faust_palm4msa2020.hpp:304-361), taken by palm4msa2
// when USE_GRAD1 is set and always on GPU2 (:633-635):
//   tmp = lambda * dynprog(L, S_j, R) - A;
//   no L: D += -lambda/c tmp R^T; else tmp = tmp R^T, then D += -lambda/c L^T tmp.
// A one-factor side is used as stored; a longer one is multiplied out with
// dynprog (the PALM product mode).
static void grad1(const Chain& pL, const Chain& pR, const DMatP& Sj, const DMat& A, double lambda, double c,
                  DMat& D, const fh_stop_crit& sc, double& error, hipStream_t st) {
  Chain lsr(pL.begin(), pL.end());
  lsr.push_back(Sj);
  lsr.insert(lsr.end(), pR.begin(), pR.end());
  DMatP tmp = lsr.size() == 1 ? dense_of(*lsr[0], st) : dynprog_multiply(lsr, st);
  const int64_t len = (int64_t)tmp->m * tmp->n;
  scale_f64(tmp->d.as<double>(), len, nullptr, lambda, 0, st);
  sub_f64(tmp->d.as<double>(), A.d.as<double>(), len, st);
  if (sc.is_criterion_error) error = fro_norm(*tmp, st);
  auto side = [&](const Chain& ch, DMatP& keep) -> const DMat* {
    if (ch.size() == 1) return ch[0].get();
    keep = dynprog_multiply(ch, st);
    return keep.get();
  };
  const double a = -lambda / c;
  DMatP Rk, Lk;
  if (!pR.empty()) {
    const DMat* R = side(pR, Rk);
    if (pL.empty()) {
      matmul_f64(*tmp, 'N', *R, 'T', a, 1.0, D, st);
      return;
    }
    auto t2 = std::make_shared<DMat>();
    matmul_f64(*tmp, 'N', *R, 'T', 1.0, 0.0, *t2, st);
    tmp = t2;
  }
  if (!pL.empty()) {
    const DMat* L = side(pL, Lk);
    matmul_f64(*L, 'T', *tmp, 'N', a, 1.0, D, st);
  }
}

// update_lambda (faust_palm4msa2020.hpp:442-493): lambda = tr(A^T S)/||S||_F^2
static void update_lambda(const Chain& S, const std::vector<Chain>& pL, const std::vector<Chain>& pR, const DMat& A,
                          double& lambda, hipStream_t st) {
  const size_t n = S.size();
  const bool packing = pR[0].size() == 1 && pL[n - 1].size() == 1;
  DMatP Smat;
  if (packing) {
    const int64_t lhs = (int64_t)S[0]->n * pR[0].front()->m;
    const int64_t rhs = (int64_t)pL[n - 1].back()->n * S[n - 1]->m;
    Chain c;
    if (lhs < rhs) {
      c.push_back(S[0]);
      for (auto& x : pR[0]) c.push_back(x);
    } else {
      c = pL[n - 1];
      c.push_back(S[n - 1]);
    }
    Smat = faust_get_product_f64(c, st);
  } else
    Smat = faust_get_product_f64(S, st);
  const int m = A.m, nc = A.n;
  DBuf diag(sizeof(double) * std::max(nc, 1)), tr(sizeof(double)), ss(sizeof(double));
  diag_dot_f64(A.d.as<double>(), Smat->d.as<double>(), m, nc, diag.as<double>(), st);
  reduce_sum_f64(diag.as<double>(), nc, tr.as<double>(), st);
  reduce_dot_f64(Smat->d.as<double>(), Smat->d.as<double>(), (int64_t)m * nc, ss.as<double>(), st);
  const double htr = read_scalar(tr, st);
  const double nS = std::sqrt(read_scalar(ss, st));
  if (nS == 0.0) throw Error(FH_ERR_NUMERIC, "Error in update_lambda: S Frobenius norm is zero, can't compute lambda.");
  if (std::isnan(htr) || std::isnan(nS))
    throw Error(FH_ERR_NUMERIC, "Error in update_lambda: S (the Faust) contains nan elements");
  lambda = htr / (nS * nS);
}

static double calc_rel_err(const Chain& S, const DMat& A, double lambda, hipStream_t st) {
  DMatP P = faust_get_product_f64(S, st);
  const int64_t N = (int64_t)A.m * A.n;
  scale_f64(P->d.as<double>(), N, nullptr, lambda, 0, st);
  sub_f64(P->d.as<double>(), A.d.as<double>(), N, st);
  DBuf s1(sizeof(double)), s2(sizeof(double));
  reduce_dot_f64(P->d.as<double>(), P->d.as<double>(), N, s1.as<double>(), st);
  reduce_dot_f64(A.d.as<double>(), A.d.as<double>(), N, s2.as<double>(), st);
  return std::sqrt(read_scalar(s1, st)) / std::sqrt(read_scalar(s2, st));
}

static int64_t sparse_size(int64_t nnz, int64_t nrows) { return nnz * (8 + 4) + (nrows + 1) * 4; }
static int64_t dense_size(int64_t nrows, int64_t ncols) { return nrows * ncols * 8; }

struct DevCons {
  fh_constraint c;
  DBuf supp;  // device copy of the support matrix (FH_PROX_SUPP)
};

// ConstraintInt / ConstraintMat / ConstraintFPP::project on a dense device matrix
static void apply_cons(const fh_constraint& c, const double* dsupp, double* M, hipStream_t st) {
  if (c.name == FH_PROX_NORMCOL || c.name == FH_PROX_NORMLIN)
    prox_real_f64(c.name, M, c.nrows, c.ncols, c.s, c.normalized != 0, c.pos != 0, st);
  else
    prox_f64(c.name, M, c.nrows, c.ncols, c.k, c.normalized != 0, c.pos != 0, dsupp, st);
}

// ConstraintInt/ConstraintMat/ConstraintFPP::project_gen (faust_prox_gen.hpp:1-25,118-210)
static DMatP project_gen(const fh_constraint& c, const double* dsupp, DMatP D, int factor_format, hipStream_t st) {
  const int64_t m = c.nrows, n = c.ncols;
  bool out_dense = false;
  switch (c.name) {
    case FH_PROX_SP: out_dense = sparse_size(c.k, m) > dense_size(m, n); break;
    case FH_PROX_SPCOL:
    case FH_PROX_SPLINCOL: out_dense = sparse_size(c.k * m, m) > dense_size(m, n); break;
    case FH_PROX_SPLIN: out_dense = sparse_size(c.k * n, m) > dense_size(m, n); break;
    default: break;
  }
  apply_cons(c, dsupp, D->d.as<double>(), st);
  D->nnz = -1;
  if (factor_format == 2) {
    // supp, normcol, normlin: the format follows the real nonzero count
    if (c.name == FH_PROX_SUPP || c.name == FH_PROX_NORMCOL || c.name == FH_PROX_NORMLIN)
      out_dense = sparse_size(nnz_of(*D, st), m) > dense_size(m, n);
    if (out_dense) return D;
    return to_sparse(*D, st);
  }
  if (factor_format == 1) return to_sparse(*D, st);
  return D;
}

struct Palm {
  const DMat* A;
  std::vector<const DevCons*> cons;
  fh_stop_crit sc;
  fh_palm_params p;
  int call_id;
  std::vector<TrajPoint>* traj;
};

static void palm4msa2(Palm& P, Chain& S, double& lambda, hipStream_t st) {
  const DMat& A = *P.A;
  const fh_palm_params& p = P.p;
  const int nf = (int)P.cons.size();
  FH_CHECK(nf > 0, "No constraint passed to palm4msa.");
  std::vector<fh_constraint> cons;
  for (auto* c : P.cons) {
    cons.push_back(c->c);
    if (p.no_normalization) cons.back().normalized = 0;
  }
  const int AllDense = 0, AllSparse = 1;
  if ((int)S.size() != nf) {
    S.clear();
    const fh_constraint& zc = p.is_update_way_R2L ? cons.back() : cons.front();
    DMatP zero = (p.factor_format != AllDense) ? dmat_zero_sparse(zc.nrows, zc.ncols, st)
                                               : dmat_zero_dense(zc.nrows, zc.ncols, st);
    if (p.is_update_way_R2L) {
      for (int i = 0; i < nf - 1; ++i) S.push_back(dmat_eye(cons[i].nrows, cons[i].ncols, p.factor_format != AllDense, st));
      S.push_back(zero);
    } else {
      S.push_back(zero);
      for (int i = 1; i < nf; ++i) S.push_back(dmat_eye(cons[i].nrows, cons[i].ncols, p.factor_format != AllDense, st));
    }
  } else if (p.factor_format == AllSparse) {
    for (auto& x : S)
      if (x->kind != K_CSR) x = to_sparse(*x, st);
  } else if (p.factor_format == AllDense) {
    for (auto& x : S)
      if (x->kind != K_DENSE) x = dense_of(*x, st);
  }
  std::vector<Chain> pL(nf), pR(nf);
  auto pack = [&](const Chain& c) { return p.packing_RL ? packed(c, st) : c; };
  double error = -1;
  int it = 0;
  int f_id = 0;
  auto do_continue = [&]() {
    if (!P.sc.is_criterion_error) return it < P.sc.num_its;
    if (error >= 0) {
      if (error < P.sc.error_threshold) return false;
      return it < P.sc.max_num_its;
    }
    return true;
  };
  while (do_continue()) {
    if (p.is_update_way_R2L) {
      pL[0].clear();
      for (int i = 1; i < nf; ++i) {
        Chain c = pL[i - 1];
        c.push_back(S[i - 1]);
        pL[i] = pack(c);
      }
      f_id = nf - 1;
    } else {
      pR[nf - 1].clear();
      for (int i = nf - 2; i >= 0; --i) {
        Chain c{S[i + 1]};
        for (auto& x : pR[i + 1]) c.push_back(x);
        pR[i] = pack(c);
      }
      f_id = 0;
    }
    while (p.is_update_way_R2L ? f_id >= 0 : f_id < nf) {
      double nR = 1, nL = 1, c;
      int flag = 0;
      if (p.constant_step_size)
        c = 1 / p.step_size;
      else {
        if (!pR[f_id].empty() && !pL[f_id].empty())
          faust_spectral_norm_pair_f64(pR[f_id], pL[f_id], p.norm2_max_iter, p.norm2_threshold, &flag, &nR, &nL, st);
        else if (!pR[f_id].empty())
          nR = faust_spectral_norm_f64(pR[f_id], p.norm2_max_iter, p.norm2_threshold, &flag, st);
        else if (!pL[f_id].empty())
          nL = faust_spectral_norm_f64(pL[f_id], p.norm2_max_iter, p.norm2_threshold, &flag, st);
        if (std::isnan(nL) || std::isnan(nR))
          throw Error(FH_ERR_NUMERIC, "2-norm computation error: R or L 2-norm is NaN.");
        if ((nR == 0 && (it > 0 || !p.is_update_way_R2L)) || (nL == 0 && (it > 0 || p.is_update_way_R2L)))
          throw Error(FH_ERR_NUMERIC,
                      "2-norm computation error in computation of descent step inverse: R or L 2-norm is zero.");
        c = 1.001 * lambda * lambda * nR * nR * nL * nL;
      }
      DMatP D = dense_of(*S[f_id], st);
      if (p.use_grad1)
        grad1(pL[f_id], pR[f_id], S[f_id], A, lambda, c, *D, P.sc, error, st);
      else
        grad2(pL[f_id], pR[f_id], A, lambda, c, *S[f_id], *D, P.sc, error, st);
      DMatP upd = project_gen(cons[f_id], P.cons[f_id]->supp.as<double>(), D, p.factor_format, st);
      if (p.factor_format == 2)
        S[f_id] = upd;  // AllDynamic: S.replace, the partial products keep the old object
      else
        // AllDense / AllSparse: S.update (faust_TransformHelper.hpp:848-868)
        // overwrites the factor object, so every partial product holding it
        // (unpacked chains, a packed chain of one dense factor) sees the update
        *S[f_id] = std::move(*upd);
      if (p.is_update_way_R2L) {
        if (f_id > 0) {
          Chain c2{S[f_id]};
          for (auto& x : pR[f_id]) c2.push_back(x);
          pR[f_id - 1] = pack(c2);
        }
        f_id--;
      } else {
        if (f_id < nf - 1) {
          Chain c2 = pL[f_id];
          c2.push_back(S[f_id]);
          pL[f_id + 1] = pack(c2);
        }
        f_id++;
      }
    }
    if (!p.no_lambda) update_lambda(S, pL, pR, A, lambda, st);
    if (P.traj) P.traj->push_back({(double)P.call_id, (double)it, calc_rel_err(S, A, lambda, st), lambda});
    it++;
  }
}

// ------------------------------------------------------------ PALM4MSA, 2016 backend
// Faust::Palm4MSA / Faust::HierarchicalFact (faust_Palm4MSA.hpp,
// faust_HierarchicalFact.hpp).  Factors are dense and carry the MatDense
// flags that change results: setZeros flags zero, setEyes flags a square
// matrix identity (faust_MatDense.hpp:243-265); gemm with a flagged operand
// copies the other operand's flags (beta = 0) or flags the result zero
// (faust_linear_algebra.hpp:600-700); sub() and the projectors' writes clear
// them.  With beta in {0, -1} the shortcut values equal the full product, so
// every product here is the canonical one; the flags decide the spectral
// norms only: 0 for a flagged-zero matrix, 1 for a flagged identity
// (MatDense::spectralNorm, faust_MatDense.hpp:660-700).
struct DF {
  DMatP M;
  bool zeros = false, ident = false;
};

static DMatP dense_new(int m, int n, hipStream_t st) {
  return make_dense(FH_F64, m, n, nullptr, std::max(m, 1), false, st);
}

static DF df_zeros(int m, int n, hipStream_t st) {
  DF d;
  d.M = dense_new(m, n, st);
  fill_f64(d.M->d.as<double>(), (int64_t)m * n, 0.0, st);
  d.zeros = true;
  return d;
}

static DF df_eye(int m, int n, hipStream_t st) {
  DF d;
  d.M = dense_new(m, n, st);
  eye_f64(d.M->d.as<double>(), m, n, st);
  d.ident = (m == n);
  return d;
}

static DF df_copy(const DF& a, hipStream_t st) {
  DF d{dense_of(*a.M, st), a.zeros, a.ident};
  return d;
}

// C = alpha op(A) op(B) + beta C_old (gemm), with the flags of the result
static DF gemm16(const DF& A, char opA, const DF& B, char opB, double alpha, double beta, const DMat* C_old,
                 hipStream_t st) {
  DF C;
  C.M = std::make_shared<DMat>();
  if (beta != 0.0) C.M = dense_of(*C_old, st);
  matmul_f64(*A.M, opA, *B.M, opB, alpha, beta, *C.M, st);
  if (beta == 0.0) {
    if (A.zeros || B.zeros) {
      fill_f64(C.M->d.as<double>(), (int64_t)C.M->m * C.M->n, 0.0, st);  // memset: +0 everywhere
      C.zeros = true;
    } else if (A.ident || B.ident) {
      const DF& other = A.ident ? B : A;
      FH_CHECK(!(other.ident && alpha != 1.0), "palm4msa2016: scaled copy of a flagged identity");
      C.zeros = other.zeros;
      C.ident = other.ident;
    }
  }
  return C;
}

static DF mul16(const DF& A, const DF& B, hipStream_t st) { return gemm16(A, 'N', B, 'N', 1.0, 0.0, nullptr, st); }

static double specnorm16(const DF& A, int maxit, double thr, hipStream_t st) {
  if (A.zeros) return 0.0;
  if (A.ident) return 1.0;
  const bool wide = A.M->m <= A.M->n;
  DF AtA = wide ? gemm16(A, 'N', A, 'T', 1.0, 0.0, nullptr, st) : gemm16(A, 'T', A, 'N', 1.0, 0.0, nullptr, st);
  int flag = 0;
  const double lam = power_iteration_f64({ChainOp{AtA.M.get(), 'N'}}, AtA.M->m, maxit, thr, &flag, st);
  return std::sqrt(std::fabs(lam));
}

// trace(X^T Y): diagonal fma chains (k ascending), canonical sum
static double trace_xty(const DMat& X, const DMat& Y, hipStream_t st) {
  DBuf dg(sizeof(double) * std::max(X.n, 1)), s(sizeof(double));
  diag_dot_f64(X.d.as<double>(), Y.d.as<double>(), X.m, X.n, dg.as<double>(), st);
  reduce_sum_f64(dg.as<double>(), X.n, s.as<double>(), st);
  return read_scalar(s, st);
}

struct Palm16 {
  DMatP data;
  std::vector<const DevCons*> cons;
  fh_stop_crit sc;
  fh_palm_params p;
  double lambda = 1.0;
  std::vector<DF> S;
  int call_id = 0;
  std::vector<TrajPoint>* traj = nullptr;

  // init_fact (faust_Palm4MSA.hpp:628-665): the first factor updated is
  // zero, the others identity
  void init_fact(int nf, hipStream_t st) {
    S.clear();
    for (int i = 0; i < nf; ++i) {
      const bool first = p.is_update_way_R2L ? (i == nf - 1) : (i == 0);
      const fh_constraint& c = cons[i]->c;
      S.push_back(first ? df_zeros(c.nrows, c.ncols, st) : df_eye(c.nrows, c.ncols, st));
    }
  }

  // compute_grad_over_c_int_opt (faust_Palm4MSA.hpp:229-330): one of four
  // associations by a flop estimate
  DMatP grad_int(const DF& L, const DF& Sj, const DF& R, double c, hipStream_t st) {
    const unsigned long long L1 = L.M->m, L2 = L.M->n, R2 = R.M->n, S2 = Sj.M->n;
    const unsigned long long cx[4] = {L1 * L2 * S2 + L1 * S2 * R2 + L2 * L1 * R2 + L2 * R2 * R2,
                                      L1 * L2 * S2 + L1 * S2 * R2 + L1 * R2 * S2 + L2 * L1 * S2,
                                      L2 * S2 * R2 + L1 * L2 * R2 + L2 * L1 * R2 + L2 * R2 * S2,
                                      L2 * S2 * R2 + L1 * L2 * R2 + L1 * R2 * S2 + L2 * L1 * S2};
    int idx = 0;
    for (int t = 1; t < 4; ++t)
      if (cx[t] < cx[idx]) idx = t;
    const bool disabled = p.grad_calc_opt_mode == 0;
    DF err;
    if (idx == 0 || idx == 1 || disabled) {
      DF t1 = mul16(L, Sj, st);
      err = gemm16(t1, 'N', R, 'N', lambda, -1.0, data.get(), st);
    } else {
      DF t1 = mul16(Sj, R, st);
      err = gemm16(L, 'N', t1, 'N', lambda, -1.0, data.get(), st);
    }
    DF g;
    if (idx == 0 || idx == 2 || disabled) {
      DF t3 = gemm16(L, 'T', err, 'N', lambda, 0.0, nullptr, st);
      g = gemm16(t3, 'N', R, 'T', 1.0 / c, 0.0, nullptr, st);
    } else {
      DF t3 = gemm16(err, 'N', R, 'T', lambda, 0.0, nullptr, st);
      g = gemm16(L, 'T', t3, 'N', 1.0 / c, 0.0, nullptr, st);
    }
    return g.M;
  }

  // compute_grad_over_c_ext_opt (faust_Palm4MSA.hpp:205-227)
  DMatP grad_ext(const DF& L, const DF& Sj, const DF& R, double c, hipStream_t st) {
    DMatP err = dense_of(*data, st);
    mul_all_ends({L.M.get(), Sj.M.get(), R.M.get()}, *err, lambda, -1.0, {'N'}, st);
    auto g = std::make_shared<DMat>();
    mul_all_ends({L.M.get(), err.get(), R.M.get()}, *g, lambda / c, 0.0, {'T', 'N', 'T'}, st);
    return g;
  }

  // next_step (faust_Palm4MSA.hpp:690-740)
  void next_step(int it, hipStream_t st) {
    const int J = (int)S.size();
    const bool R2L = p.is_update_way_R2L;
    std::vector<DF> RorL(J);
    DF LorR;
    if (!R2L) {
      LorR = df_eye(cons[0]->c.nrows, cons[0]->c.nrows, st);
      RorL[J - 1] = df_eye(cons[J - 1]->c.ncols, cons[J - 1]->c.ncols, st);
      for (int i = J - 2; i >= 0; --i) RorL[i] = mul16(S[i + 1], RorL[i + 1], st);
    } else {
      LorR = df_eye(cons[J - 1]->c.ncols, cons[J - 1]->c.ncols, st);
      RorL[0] = df_eye(cons[0]->c.nrows, cons[0]->c.nrows, st);
      for (int i = 0; i < J - 1; ++i) RorL[i + 1] = mul16(RorL[i], S[i], st);
    }
    for (int j = 0; j < J; ++j) {
      const int f = R2L ? J - 1 - j : j;
      double c = 1.0 / p.step_size;
      if (!p.constant_step_size) {  // compute_c (faust_Palm4MSA.hpp:612-625)
        const double nL = specnorm16(LorR, p.norm2_max_iter, p.norm2_threshold, st);
        const double nR = specnorm16(RorL[f], p.norm2_max_iter, p.norm2_threshold, st);
        c = 1.001 * nR * nR * nL * nL * lambda * lambda;
      }
      const DF& L = R2L ? RorL[f] : LorR;
      const DF& R = R2L ? LorR : RorL[f];
      DMatP g = p.grad_calc_opt_mode == 2 ? grad_ext(L, S[f], R, c, st) : grad_int(L, S[f], R, c, st);
      // compute_projection: S -= grad/c, then project (the flags are cleared)
      const fh_constraint& cc = cons[f]->c;
      sub_f64(S[f].M->d.as<double>(), g->d.as<double>(), (int64_t)cc.nrows * cc.ncols, st);
      S[f].zeros = S[f].ident = false;
      apply_cons(cc, cons[f]->supp.as<double>(), S[f].M->d.as<double>(), st);
      S[f].M->nnz = -1;
      LorR = R2L ? mul16(S[f], LorR, st) : mul16(LorR, S[f], st);
    }
    // compute_lambda (faust_Palm4MSA.hpp:333-358)
    const double tr_hh = trace_xty(*LorR.M, *LorR.M, st);
    if (tr_hh == 0.0) throw Error(FH_ERR_NUMERIC, "compute_lambda : Xhatt_Xhat_tr equal 0 so lambda is infinite");
    lambda = std::fabs(trace_xty(*data, *LorR.M, st) / tr_hh);
    if (std::isnan(lambda)) throw Error(FH_ERR_NUMERIC, "compute_lambda : lambda is NaN");
    if (traj) traj->push_back({(double)call_id, (double)it, std::nan(""), lambda});
  }

  // compute_facts (faust_Palm4MSA.hpp:170-183), StoppingCriterion counting
  void compute_facts(hipStream_t st) {
    if (sc.is_criterion_error)
      throw Error(FH_ERR_UNSUPPORTED, "2016 backend: only the iteration-count stopping criterion");
    for (int it = 0; it < sc.num_its; ++it) next_step(it, st);
  }
};

static std::vector<std::unique_ptr<DevCons>> upload_cons(const fh_constraint* c, int n, hipStream_t st) {
  std::vector<std::unique_ptr<DevCons>> out;
  for (int i = 0; i < n; ++i) {
    auto d = std::make_unique<DevCons>();
    d->c = c[i];
    FH_CHECK(c[i].name >= FH_PROX_SP && c[i].name <= FH_PROX_NORMLIN, "unknown constraint name");
    if (c[i].name == FH_PROX_SUPP) {
      FH_CHECK(c[i].supp != nullptr, "supp constraint needs its support matrix");
      size_t b = sizeof(double) * (size_t)c[i].nrows * c[i].ncols;
      d->supp.alloc(b);
      FH_HIP(hipMemcpyAsync(d->supp.p, c[i].supp, b, hipMemcpyHostToDevice, st));
    }
    out.push_back(std::move(d));
  }
  FH_HIP(hipStreamSynchronize(st));
  return out;
}

}  // namespace fh

struct fh_palm_result_s {
  fh::Chain S;
  double lambda = 1.0;
  std::vector<fh::TrajPoint> traj;
};

using namespace fh;

extern "C" {

int fh_palm4msa(const double* A, int nrows, int ncols, const fh_constraint* cons, int ncons, const fh_stop_crit* sc,
                const fh_palm_params* p, double lambda_init, fh_faust_t init_facts, fh_palm_result_t* out) {
  FH_API_BEGIN
  FH_CHECK(A && cons && sc && p && out, "null argument");
  FH_CHECK(ncons > 0, "No constraint passed to palm4msa.");
  hipStream_t st = nullptr;
  auto r = std::make_unique<fh_palm_result_s>();
  DMatP Ad = make_dense(FH_F64, nrows, ncols, A, std::max(nrows, 1), false, st);
  auto dc = upload_cons(cons, ncons, st);
  Palm P{Ad.get(), {}, *sc, *p, 0, p->record_traj ? &r->traj : nullptr};
  for (auto& c : dc) P.cons.push_back(c.get());
  if (init_facts) {
    for (auto& f : init_facts->f) {
      FH_CHECK(f->dtype == FH_F64 && f->kind != K_BSR, "init factors must be FP64 dense or CSR");
      r->S.push_back(f);
    }
  }
  r->lambda = lambda_init;
  palm4msa2(P, r->S, r->lambda, st);
  FH_HIP(hipStreamSynchronize(st));
  *out = r.release();
  FH_API_END
}

int fh_hierarchical(const double* A, int nrows, int ncols, const fh_constraint* fac_cons,
                    const fh_constraint* res_cons, int nlevels, const fh_stop_crit* sc_local,
                    const fh_stop_crit* sc_global, const fh_palm_params* p, fh_palm_result_t* out) {
  FH_API_BEGIN
  FH_CHECK(A && fac_cons && res_cons && sc_local && sc_global && p && out, "null argument");
  hipStream_t st = nullptr;
  auto r = std::make_unique<fh_palm_result_s>();
  DMatP Ad = make_dense(FH_F64, nrows, ncols, A, std::max(nrows, 1), false, st);
  auto fc = upload_cons(fac_cons, nlevels, st);
  auto rc = upload_cons(res_cons, nlevels, st);
  Chain& S = r->S;
  S.push_back(dense_of(*Ad, st));
  double glo_lambda = 1.0;
  const bool R2L = p->is_update_way_R2L, left = p->is_fact_side_left;
  std::vector<TrajPoint>* traj = p->record_traj ? &r->traj : nullptr;
  for (int i = 0; i < nlevels; ++i) {
    DMatP Si = left ? S.front() : S[i];
    const fh_constraint &fcc = fc[i]->c, &rcc = rc[i]->c;
    int zd[2], idd[2];
    if ((!R2L && !left) || (R2L && left)) {
      zd[0] = fcc.nrows; zd[1] = fcc.ncols; idd[0] = rcc.nrows; idd[1] = rcc.ncols;
    } else {
      zd[0] = rcc.nrows; zd[1] = rcc.ncols; idd[0] = fcc.nrows; idd[1] = fcc.ncols;
    }
    const bool dense = p->factor_format == 0;
    DMatP zero = dense ? dmat_zero_dense(zd[0], zd[1], st) : dmat_zero_sparse(zd[0], zd[1], st);
    DMatP id = dmat_eye(idd[0], idd[1], !dense, st);
    Chain Sith = R2L ? Chain{id, zero} : Chain{zero, id};
    double lambda_ = 1.0;
    DMatP Sid = dense_of(*Si, st);
    Palm P1{Sid.get(), {}, *sc_local, *p, 2 * i, traj};
    if (left) P1.cons = {rc[i].get(), fc[i].get()}; else P1.cons = {fc[i].get(), rc[i].get()};
    palm4msa2(P1, Sith, lambda_, st);
    glo_lambda *= lambda_;
    if (left) {
      S.erase(S.begin());
      S.insert(S.begin(), Sith[1]);
      S.insert(S.begin(), Sith[0]);
    } else {
      S.pop_back();
      S.push_back(Sith[0]);
      S.push_back(Sith[1]);
    }
    Palm P2{Ad.get(), {}, *sc_global, *p, 2 * i + 1, traj};
    for (int t = 0; t <= i; ++t) {
      if (left) P2.cons.insert(P2.cons.begin(), fc[t].get()); else P2.cons.push_back(fc[t].get());
    }
    if (left) P2.cons.insert(P2.cons.begin(), rc[i].get()); else P2.cons.push_back(rc[i].get());
    palm4msa2(P2, S, glo_lambda, st);
  }
  r->lambda = glo_lambda;
  FH_HIP(hipStreamSynchronize(st));
  *out = r.release();
  FH_API_END
}

int fh_palm4msa2016(const double* A, int nrows, int ncols, const fh_constraint* cons, int ncons,
                    const fh_stop_crit* sc, const fh_palm_params* p, double lambda_init, fh_faust_t init_facts,
                    fh_palm_result_t* out) {
  FH_API_BEGIN
  FH_CHECK(A && cons && sc && p && out, "null argument");
  FH_CHECK(ncons > 0, "No constraint passed to palm4msa.");
  hipStream_t st = nullptr;
  auto r = std::make_unique<fh_palm_result_s>();
  auto dc = upload_cons(cons, ncons, st);
  Palm16 P;
  P.data = make_dense(FH_F64, nrows, ncols, A, std::max(nrows, 1), false, st);
  for (auto& c : dc) P.cons.push_back(c.get());
  P.sc = *sc;
  P.p = *p;
  P.lambda = lambda_init;
  P.traj = p->record_traj ? &r->traj : nullptr;
  if (init_facts && !init_facts->f.empty()) {
    FH_CHECK((int)init_facts->f.size() == ncons, "palm4msa2016: init_facts and constraints differ in number");
    for (auto& f : init_facts->f) {
      FH_CHECK(f->dtype == FH_F64 && f->kind != K_BSR, "init factors must be FP64 dense or CSR");
      P.S.push_back(DF{dense_of(*f, st), false, false});
    }
  } else {
    P.init_fact(ncons, st);
  }
  P.compute_facts(st);
  for (auto& d : P.S) r->S.push_back(d.M);
  r->lambda = P.lambda;
  FH_HIP(hipStreamSynchronize(st));
  *out = r.release();
  FH_API_END
}

int fh_hierarchical2016(const double* A, int nrows, int ncols, const fh_constraint* fac_cons,
                        const fh_constraint* res_cons, int nlevels, const fh_stop_crit* sc_local,
                        const fh_stop_crit* sc_global, const fh_palm_params* p, double init_lambda,
                        fh_palm_result_t* out) {
  FH_API_BEGIN
  FH_CHECK(A && fac_cons && res_cons && sc_local && sc_global && p && out, "null argument");
  FH_CHECK(nlevels > 0, "hierarchical2016: no level");
  hipStream_t st = nullptr;
  auto r = std::make_unique<fh_palm_result_s>();
  auto fc = upload_cons(fac_cons, nlevels, st);
  auto rc = upload_cons(res_cons, nlevels, st);
  const bool left = p->is_fact_side_left;
  std::vector<TrajPoint>* traj = p->record_traj ? &r->traj : nullptr;
  Palm16 P2, PG;
  P2.data = make_dense(FH_F64, nrows, ncols, A, std::max(nrows, 1), false, st);
  PG.data = P2.data;
  P2.sc = *sc_local;
  PG.sc = *sc_global;
  P2.p = PG.p = *p;
  P2.traj = PG.traj = traj;
  PG.lambda = init_lambda;
  // HierarchicalFact::init / next_step (faust_HierarchicalFact.hpp:95-150)
  PG.cons = {left ? fc[0].get() : rc[0].get()};
  PG.init_fact(1, st);
  for (int i = 0; i < nlevels; ++i) {
    P2.cons = {fc[i].get(), rc[i].get()};
    P2.init_fact(2, st);
    P2.lambda = init_lambda;
    P2.call_id = 2 * i;
    P2.compute_facts(st);
    PG.lambda *= P2.lambda;  // update_lambda_from_palm
    if (left) {
      PG.cons[0] = fc[i].get();
      PG.cons.insert(PG.cons.begin() + 1, rc[i].get());
      PG.S.insert(PG.S.begin(), df_copy(P2.S[0], st));  // init_fact_from_palm
      PG.S[1] = df_copy(P2.S[1], st);
    } else {
      PG.cons.insert(PG.cons.begin() + i, fc[i].get());
      PG.cons[i + 1] = rc[i].get();
      PG.S.back() = df_copy(P2.S[0], st);
      PG.S.push_back(df_copy(P2.S[1], st));
    }
    PG.call_id = 2 * i + 1;
    PG.compute_facts(st);
    P2.data = dense_of(*(left ? PG.S[0] : PG.S[i + 1]).M, st);  // set_data(get_res(...))
  }
  for (auto& d : PG.S) r->S.push_back(d.M);
  r->lambda = PG.lambda;
  FH_HIP(hipStreamSynchronize(st));
  *out = r.release();
  FH_API_END
}

int fh_palm_result_nfacts(fh_palm_result_t r, int* n) {
  FH_API_BEGIN
  FH_CHECK(r && n, "null argument");
  *n = (int)r->S.size();
  FH_API_END
}

int fh_palm_result_lambda(fh_palm_result_t r, double* lambda) {
  FH_API_BEGIN
  FH_CHECK(r && lambda, "null argument");
  *lambda = r->lambda;
  FH_API_END
}

int fh_palm_result_factor(fh_palm_result_t r, int i, fh_mat_t* out) {
  FH_API_BEGIN
  FH_CHECK(r && out && i >= 0 && i < (int)r->S.size(), "bad factor index");
  *out = new fh_mat_s{r->S[i]};
  FH_API_END
}

int fh_palm_result_traj(fh_palm_result_t r, double* out4, int* len) {
  FH_API_BEGIN
  FH_CHECK(r && len, "null argument");
  if (out4)
    for (size_t i = 0; i < r->traj.size(); ++i) {
      out4[4 * i] = r->traj[i].call;
      out4[4 * i + 1] = r->traj[i].it;
      out4[4 * i + 2] = r->traj[i].relerr;
      out4[4 * i + 3] = r->traj[i].lambda;
    }
  *len = (int)r->traj.size();
  FH_API_END
}

int fh_palm_result_free(fh_palm_result_t r) {
  FH_API_BEGIN
  delete r;
  FH_API_END
}

}  // extern "C"
