/*
 * faust_oracle.h -- C interface of the CPU FP64 oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C++ restatement of the
 * reference FAuST algorithms on the two hot paths (Faust x dense products and
 * the PALM4MSA / hierarchical factorization loop).  It is loaded by the test
 * suite, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the
 * checker.  Nothing in the product (faust_amd/) links or calls it.
 *
 * Arithmetic convention shared with the HIP kernels (the "canonical order"):
 *   - every matrix-product entry is acc = 0; for k ascending: acc = fma(a,b,acc)
 *     (terms with an exact-zero operand are skipped, which leaves acc unchanged);
 *   - alpha/beta epilogue: t = alpha*acc; out = beta==0 ? t : fma(beta, c, t);
 *   - sums/dots over n values: 4096-element chunks, 256 lanes per chunk, lane t
 *     takes elements t, t+256, ..., a fixed halving tree over the 256 lanes,
 *     then the chunk partials are added sequentially in chunk order.
 * Eigen's own summation order cannot be reproduced without Eigen (absent);
 * this convention replaces it, and the GPU path follows it bit for bit.
 */
#ifndef FAUST_ORACLE_H
#define FAUST_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* factor kinds (MatDense / MatSparse / MatBSR) */
enum { OR_DENSE = 0, OR_SPARSE = 1, OR_BSR = 2 };

/* A factor handed across the C boundary: CSR int32 / col-major dense /
 * BSR with column-major bm x bn blocks (faust_MatBSR.h:120-150). */
typedef struct {
  int kind;
  int nrows, ncols;
  /* OR_DENSE */
  const double *dense;           /* nrows*ncols col-major */
  /* OR_SPARSE */
  int nnz;
  const int *rowptr, *colind;
  const double *vals;
  /* OR_BSR */
  int bm, bn, bnnz;
  const int *browptr, *bcolinds;
  const double *bdata;
} or_fac;

/* constraint names (subset of faust_ConstraintGeneric.h:49-70) */
enum { OR_SP = 0, OR_SPCOL = 1, OR_SPLIN = 2, OR_SPLINCOL = 3, OR_SUPP = 4, OR_NORMCOL = 5, OR_NORMLIN = 6 };

typedef struct {
  int name;
  int k;
  int nrows, ncols;
  int normalized;
  int pos;
  const double *supp; /* OR_SUPP only, nrows*ncols col-major */
  double s;           /* OR_NORMCOL / OR_NORMLIN: the 2-norm imposed on every column / row */
} or_cons;

typedef struct {
  int num_its;
  int is_criterion_error;
  double error_threshold;
  int max_num_its;
} or_stop;

typedef struct {
  int is_update_way_R2L;
  int factor_format;    /* 0 dense, 1 sparse, 2 dynamic */
  int packing_RL;
  int no_normalization;
  int no_lambda;
  int norm2_max_iter;
  double norm2_threshold;
  int constant_step_size;
  double step_size;
  int is_fact_side_left; /* hierarchical only */
  int use_grad1;         /* compute_n_apply_grad1 (USE_GRAD1, and the GPU2 path) */
  int grad_calc_opt_mode; /* 2016 backend: 0 disabled, 1 internal, 2 external (pyfaust's default) */
} or_params;

const char *or_last_error(void);

/* canonical reductions */
double or_reduce_sum(const double *x, int64_t n);
double or_reduce_dot(const double *x, const double *y, int64_t n);

/* C = alpha*op(A)*op(B) + beta*C, dense col-major operands */
int or_gemm(char opA, char opB, int m, int n, int k, double alpha,
            const double *A, int lda, const double *B, int ldb, double beta,
            double *C, int ldc);

/* Y = op(F) X, F = chain of factors (faust_Transform.hpp:1714-1812).
 * X: (op=='N' ? ncols(F) : nrows(F)) x p col-major. */
int or_chain_multiply(const or_fac *facs, int nfacs, char op, const double *X,
                      int p, double *Y, int nthreads);

/* dense product of the chain (faust_Transform.hpp:599-652, op 'N') */
int or_get_product(const or_fac *facs, int nfacs, double *out);

/* dynprog_multiply's tables (faust_prod_opt_gen.hpp:81-200): cost[i*n+j] the
 * wrapping int32 cost of the best product of factors i..j, split[i*n+j] its
 * split k, i.e. (i..k)(k+1..j) */
int or_dynprog_tables(const or_fac *facs, int nfacs, int32_t *cost, int *split);

/* spectral norm by power iteration on the F F^H / F^H F chain
 * (faust_Transform.hpp:684-709, faust_linear_algebra.hpp:808-898) */
int or_spectral_norm(const or_fac *facs, int nfacs, int maxit, double thr,
                     double *out_norm, int *out_flag);

/* the start vector power_iteration draws: srand(0xF4+'u'+57); Random(n) */
void or_power_iteration_x0(int n, double *out);

/* proximal projectors on a dense col-major matrix (faust_prox.hpp) */
int or_prox(const or_cons *c, double *M);

/* multiply_order_opt_all_ends (faust_prod_opt.hpp:49-86), the product order
 * of PALM4MSA's gradient: out = alpha * op(F_0) ... op(F_{n-1}) + beta * out
 * (out: out_m x out_n col-major, read for beta); flags as the reference's
 * transconj_flags: one per factor, or one flag that the intermediate products
 * apply to every factor and the last product to its left operand only */
int or_mul_all_ends(const or_fac *facs, int nfacs, const char *flags, int nflags,
                    double alpha, double beta, int out_m, int out_n, double *out);

/* the format ConstraintInt/ConstraintMat::project_gen returns for M
 * (faust_prox_gen.hpp: sparse_size vs dense_size): OR_DENSE or OR_SPARSE */
int or_project_gen_kind(const or_cons *c, const double *M, int *kind);

/* PALM4MSA 2020 / hierarchical results */
typedef struct or_result or_result;
int or_palm4msa(const double *A, int m, int n, const or_cons *cons, int ncons,
                const or_stop *sc, const or_params *p, double lambda_in,
                const or_fac *init_facs, int n_init, int record_traj,
                or_result **out);
int or_hierarchical(const double *A, int m, int n, const or_cons *fac_cons,
                    const or_cons *res_cons, int nlevels, const or_stop *sc_local,
                    const or_stop *sc_global, const or_params *p,
                    int record_traj, or_result **out);
/* The 2016 backend (faust_Palm4MSA.hpp, faust_HierarchicalFact.hpp): dense
 * factors; the trajectory holds lambda after each iteration (relerr NaN). */
int or_palm4msa2016(const double *A, int m, int n, const or_cons *cons, int ncons,
                    const or_stop *sc, const or_params *p, double lambda_in,
                    const or_fac *init_facs, int n_init, int record_traj,
                    or_result **out);
int or_hierarchical2016(const double *A, int m, int n, const or_cons *fac_cons,
                        const or_cons *res_cons, int nlevels, const or_stop *sc_local,
                        const or_stop *sc_global, const or_params *p,
                        double init_lambda, int record_traj, or_result **out);
int or_result_nfacts(const or_result *r);
double or_result_lambda(const or_result *r);
int or_result_fact_info(const or_result *r, int i, int *kind, int *nrows,
                        int *ncols, int *nnz);
int or_result_fact_dense(const or_result *r, int i, double *out);
int or_result_traj_len(const or_result *r);
int or_result_traj(const or_result *r, double *out);
void or_result_free(or_result *r);

#ifdef __cplusplus
}
#endif
#endif
