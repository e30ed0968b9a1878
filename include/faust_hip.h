/*
 * faust_hip.h -- C ABI of libfaust_hip.so, the MI355X (gfx950) backend of the
 * FAuST hot paths.  Plain pointers and sizes only; every entry point returns
 * an int status (FH_OK == 0) and fh_last_error() gives the message of the
 * last failure on the calling thread.
 *
 * The roles mirror the gpu_mod C function tables that the reference's GPU2
 * layer dlopens (src/faust_linear_operator/GPU2/faust_gpu_mod_utils.h:24-102,
 * faust_gpu_mod_utils.cpp:40-90): dense / sparse (CSR) / BSR matrix handles
 * with togpu/tocpu/free, `mul_gpu_dsm_ext`-style products with op flags and
 * alpha/beta passed by pointer, prox projectors, norms and device selection.
 * On top of that it exports the two hot-path drivers the reference's C++
 * core runs on the CPU: the Faust chain product (TransformHelper::multiply)
 * and palm4msa2 / hierarchical.
 *
 * Conventions (as at the reference call sites):
 *   - dense matrices are column-major with a leading dimension (ld >= nrows);
 *   - sparse matrices are CSR with int32 rowptr/colind (faust_MatSparse.h:261-264);
 *   - BSR: int32 browptr/bcolinds, blocks of bm x bn stored column-major
 *     (faust_MatBSR.h:120-150);
 *   - op: 'N', 'T' or 'H' (gm_Op OP_NOTRANSP/OP_TRANSP/OP_CONJTRANSP,
 *     faust_gpu_mod_utils.cpp:189-199); the scalars here are real, so 'H' == 'T';
 *   - `stream` is a hipStream_t (NULL = the null stream); results are ordered
 *     on that stream and the call does not wait unless stated.
 */
#ifndef FAUST_HIP_H
#define FAUST_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { FH_OK = 0, FH_ERR_ARG = 1, FH_ERR_HIP = 2, FH_ERR_NUMERIC = 3, FH_ERR_UNSUPPORTED = 4 };
enum { FH_F64 = 0, FH_F32 = 1, FH_BF16 = 2 };
enum { FH_MAT_DENSE = 0, FH_MAT_CSR = 1, FH_MAT_BSR = 2 };

typedef struct fh_mat_s *fh_mat_t;     /* device matrix handle (dense, CSR or BSR) */
typedef struct fh_faust_s *fh_faust_t; /* chain of factor handles (a Faust) */

/* Message of the last failed call on this thread (never NULL). */
const char *fh_last_error(void);
/* Library version string. */
const char *fh_version(void);

/* Forget the execution forms the autotuner measured (in FH_MUL_AUTO mode the
 * first product of each chain signature outside graph capture times the
 * forms that give the same bits -- FP64 CSR chains: one launch per factor and
 * the one-launch chains with their tile choices; BSR chains: kernels 5, 1, 2,
 * 4, 0; FP32 CSR chains: one or two rows per thread -- and keeps the
 * quickest; FH_AUTOTUNE=0 turns it off).  No counterpart in the reference,
 * whose optimize_multiply is an explicit call (faust_TransformHelper.hpp:436-512). */
int fh_autotune_clear(void);

/* ---- device selection: gm_GenPurposeFunc dev_count/set_dev/cur_dev
 *      (src/faust_linear_operator/GPU2/faust_cuda_device.cpp:6-30) */
int fh_dev_count(int *count);
int fh_set_dev(int dev);
int fh_cur_dev(int *dev);
/* Wait for all work on `stream`. */
int fh_stream_sync(void *stream);

/* ---- matrix handles: togpu / tocpu / info / free
 *      (dsm_funcs->togpu, spm_funcs->togpu, bsr_funcs->togpu;
 *       faust_MatDense_gpu.cpp.in:115, faust_MatSparse_gpu.cpp.in, faust_MatBSR_gpu.cpp.in).
 * `src_on_device` != 0 means the pointers are device pointers (copied D2D). */
int fh_mat_dense_create(int dtype, int nrows, int ncols, const void *data, int ld, int src_on_device,
                        fh_mat_t *out);
int fh_mat_csr_create(int dtype, int nrows, int ncols, int nnz, const int32_t *rowptr, const int32_t *colind,
                      const void *vals, int src_on_device, fh_mat_t *out);
int fh_mat_bsr_create(int dtype, int nrows, int ncols, int bm, int bn, int bnnz, const int32_t *browptr,
                      const int32_t *bcolinds, const void *bdata, int src_on_device, fh_mat_t *out);
int fh_mat_info(fh_mat_t M, int *kind, int *dtype, int *nrows, int *ncols, int64_t *nnz);
/* Copy back to host buffers (dense: `data` nrows*ncols col-major; CSR:
 * rowptr[nrows+1], colind[nnz], vals[nnz]; BSR (32x32 blocks): browptr
 * [nrows/32+1], bcolinds[bnnz], bdata[bnnz*1024] with the blocks column-major
 * one after the other, as BF16 bit patterns for BF16 handles and as FP32 for
 * FP32 handles, whose uploaded values come back exactly).  For BSR handles
 * `nnz` of fh_mat_info is bnnz*1024.  Any output pointer may be NULL to skip it. */
int fh_mat_tocpu(fh_mat_t M, void *data_or_vals, int32_t *rowptr, int32_t *colind);
/* Densify into a device buffer (col-major, ld = nrows): FP64 for FP64
 * matrices, FP32 for FP32 matrices and for BSR matrices of either value type,
 * BF16 for BF16 dense matrices. */
int fh_mat_to_dense(fh_mat_t M, void *dense_dev, void *stream);
int fh_mat_free(fh_mat_t M);

/* ---- products
 * Y = alpha*op(M)*X + beta*Y with M a matrix handle and X, Y device dense
 * col-major.  Replaces spm_funcs->mul_gpu_dsm_ext / bsr_funcs->mul_gpu_dsm_ext
 * / dsm_funcs->mul_gpu_dsm_ext (faust_MatSparse_gpu.cpp.in:456-472,
 * faust_MatBSR_gpu.cpp.in:109-152, faust_MatDense_gpu.cpp.in:478).
 * alpha/beta point to host scalars of the matrix dtype (FP32 for BF16). */
int fh_mat_mul_dense(fh_mat_t M, char op, const void *X, int ldx, int ncols, void *Y, int ldy,
                     const void *alpha, const void *beta, void *stream);
/* C = alpha*op(A)*op(B) + beta*C, dense device operands (dsm_funcs->gemm,
 * faust_MatDense_gpu.cpp.in:518-551; FP64 follows the canonical order). */
int fh_gemm(int dtype, char opA, char opB, int m, int n, int k, const void *alpha, const void *A, int lda,
            const void *B, int ldb, const void *beta, void *C, int ldc, void *stream);

/* ---- Faust (factor chain).  Handles are borrowed: freeing the Faust does
 * not free the factors (faust_TransformHelperGen.h:140 shares factors). */
int fh_faust_create(fh_faust_t *out);
int fh_faust_push_back(fh_faust_t F, fh_mat_t M);
int fh_faust_size(fh_faust_t F, int *nfactors);
/* factor i of F as a new matrix handle owned by the caller; the factor is
 * shared, not copied (TransformHelper::get_gen_fact,
 * faust_TransformHelperGen.h:63, the getter behind pyfaust's Faust.factors). */
int fh_faust_factor(fh_faust_t F, int i, fh_mat_t *out);
int fh_faust_free(fh_faust_t F);
/* Y = op(F)*X, X (ncols(op(F)) x p) and Y device col-major with ld = rows.
 * TransformHelper<FPP,Cpu>::multiply(const FPP*, int, FPP*)
 * (faust_TransformHelper.hpp:407-434) -> Transform::multiply
 * (faust_Transform.hpp:1714-1812): right-to-left for 'N', left-to-right with
 * per-factor op for 'T'/'H', two ping-pong intermediates.  p = 0 is a no-op
 * (X and Y may then be null). */
int fh_faust_multiply(fh_faust_t F, char op, const void *X, int p, void *Y, void *stream);
/* fh_faust_multiply without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The two ping-pong intermediates live in a fixed per-GPU device array of
 * FH_NOALLOC_SCRATCH_BYTES (held from library load to process end) instead of
 * the Faust's own buffers; the kernels, and so the results, are those of
 * fh_faust_multiply.  A product whose intermediates need more than the array
 * (each rounded up to 256 bytes) is refused with FH_ERR_ARG.  FP32 chains with
 * dense FP32 factors need their CSR twins built by one fh_faust_multiply call
 * first (FH_ERR_UNSUPPORTED otherwise).  Per-factor timers are not recorded.
 * Every call on one GPU shares the array: calls that overlap on one GPU (other
 * streams, other threads) race and must be serialized by the caller. */
#define FH_NOALLOC_SCRATCH_BYTES (32u << 20)
int fh_faust_multiply_noalloc(fh_faust_t F, char op, const void *X, int p, void *Y, void *stream);
/* Number of product kernels the last fh_faust_multiply on F launched (one
 * per factor, plus the input transpose of FP32 CSR chains).  No counterpart
 * in the reference; for benchmarks and tests. */
int fh_faust_last_launches(fh_faust_t F, int *launches);
/* Execution form of F's products (TransformHelper::set_FM_mul_mode,
 * faust_TransformHelper.hpp:436-512 optimize_multiply, which times the modes
 * and keeps the quickest; faust_amd.Faust.optimize_time does the same on the
 * GPU): FH_MUL_AUTO (defaults; for FP64 CSR chains the autotuned form, see
 * fh_autotune_clear), FH_MUL_PER_FACTOR / FH_MUL_FUSED (FP64 CSR chains: one
 * launch per factor / the staged one-launch chain when it fits LDS),
 * FH_MUL_BSR_KERNEL(k) (BSR chains: kernel k = 0..5, as FH_BSR_KERNEL),
 * FH_MUL_CSR_ROWS1 / FH_MUL_CSR_ROWS2 (FP32 CSR chains: rows per thread).
 * A mode that does not apply to F's chain kind falls back to the default. */
#define FH_MUL_AUTO 0
#define FH_MUL_PER_FACTOR 1
#define FH_MUL_FUSED 2
#define FH_MUL_BSR_KERNEL(k) (16 + (k))
#define FH_MUL_CSR_ROWS1 32
#define FH_MUL_CSR_ROWS2 33
int fh_faust_set_mul_mode(fh_faust_t F, int mode);
/* Activations of a BSR chain (no counterpart needed in the reference, whose
 * MatBSR<float> products are FP32 throughout, faust_MatBSR.hpp:828-836):
 * FH_BF16 (the default: BF16 X / Y and intermediates, FP32 accumulation on
 * BF16 MFMA, FP32 values as a BF16 hi/lo pair) or FH_F32 (FP32 X / Y and
 * intermediates, FP32 blocks on v_mfma_f32_32x32x2_f32: the reference's
 * precision), the latter for chains of FP32 BSR factors only
 * (FH_ERR_UNSUPPORTED otherwise).  fh_faust_get_act_dtype gives the dtype X
 * and Y have in fh_faust_multiply for any chain. */
int fh_faust_set_act_dtype(fh_faust_t F, int dtype);
int fh_faust_get_act_dtype(fh_faust_t F, int *dtype);
int fh_faust_get_mul_mode(fh_faust_t F, int *mode);
/* Per-factor device timers: HIP events recorded on the launch stream around
 * each factor's kernel of every fh_faust_multiply while enabled (the
 * reference's per-factor Transform::multiply timers, faust_Transform.hpp:
 * 1499-1536, behind __COMPILE_TIMERS__).  fh_faust_timers fills ms[i] and
 * counts[i] for factor i (i < nfactors) and slot nfactors for layout
 * conversions, then resets them; arrays hold nfactors+1 entries.  Synchronizes. */
int fh_faust_enable_timers(fh_faust_t F, int on);
int fh_faust_timers(fh_faust_t F, double *ms, int64_t *counts);
/* Dense FP64 product of an FP64 dense / CSR chain into a device buffer
 * (Transform::get_product, faust_Transform.hpp:599-652); other chains give
 * FH_ERR_UNSUPPORTED (their product is F times the identity). */
int fh_faust_get_product(fh_faust_t F, void *out_dev, void *stream);
/* 2-norm by power iteration (TransformHelper::spectralNorm,
 * faust_TransformHelper.hpp:1201-1218; power_iteration,
 * faust_linear_algebra.hpp:808-898) of an FP64 dense / CSR chain; other
 * chains give FH_ERR_UNSUPPORTED.  Synchronizes. */
int fh_faust_spectral_norm(fh_faust_t F, int max_iter, double threshold, double *norm, int *flag, void *stream);

/* ---- projectors on a dense device matrix, in place
 * (dsm_funcs->prox_sp/prox_spcol/prox_splin, faust_MatDense_gpu.cpp.in:746-765;
 *  splincol and supp are native here, the reference falls back to the CPU:
 *  faust_prox_gpu.hpp:103-108).  Tie handling reproduces std::partial_sort
 *  with comparator |a| > |b| (faust_prox.hpp:53-69). */
enum { FH_PROX_SP = 0, FH_PROX_SPCOL = 1, FH_PROX_SPLIN = 2, FH_PROX_SPLINCOL = 3, FH_PROX_SUPP = 4,
       FH_PROX_NORMCOL = 5, FH_PROX_NORMLIN = 6 };
int fh_prox(int dtype, int name, void *M, int nrows, int ncols, int64_t k, int normalized, int pos,
            const void *supp, void *stream);
/* normcol / normlin (Faust::prox_normcol / prox_normlin, faust_prox.hpp:205-255;
 * the reference's GPU2 form runs them on the CPU, faust_prox_gpu.hpp:137-150):
 * every column (row) scaled to 2-norm s, zero lines left zero, s == 0 zeroes
 * the matrix, s < 0 is an FH_ERR_ARG; then the optional Frobenius
 * normalisation.  FP64 only. */
int fh_prox_real(int dtype, int name, void *M, int nrows, int ncols, double s, int normalized, int pos,
                 void *stream);

/* ---- canonical FP64 reductions (dsm_funcs->norm_frob / dot); synchronize. */
int fh_norm_frob(int dtype, const void *M, int64_t n, double *out, void *stream);
int fh_dot(int dtype, const void *x, const void *y, int64_t n, double *out, void *stream);

/* ---- PALM4MSA 2020 / hierarchical (faust_palm4msa2020.hpp:47-300,
 * faust_hierarchical.hpp:1-150; Python entry FaustFact.hpp:580-675,740-860). */
typedef struct {
  int name;          /* FH_PROX_* */
  int64_t k;
  int nrows, ncols;
  int normalized;    /* ConstraintInt default: 1; ConstraintReal (normcol/normlin): 0 */
  int pos;           /* default: 0 */
  const double *supp;/* FH_PROX_SUPP: host nrows*ncols col-major */
  double s;          /* FH_PROX_NORMCOL / FH_PROX_NORMLIN: the 2-norm of every column / row */
} fh_constraint;

typedef struct {
  int num_its;            /* StoppingCriterion num_its (factparams.py:20-31) */
  int is_criterion_error; /* stop on error < error_threshold */
  double error_threshold;
  int max_num_its;
} fh_stop_crit;

typedef struct {
  int is_update_way_R2L;  /* factparams.py:955-962 defaults */
  int factor_format;      /* 0 dense, 1 sparse, 2 dynamic */
  int packing_RL;
  int no_normalization;
  int no_lambda;
  int norm2_max_iter;
  double norm2_threshold;
  int constant_step_size;
  double step_size;
  int is_fact_side_left;  /* hierarchical */
  int record_traj;        /* record (call, it, rel_err, lambda) after each iteration */
  int use_grad1;          /* gradient as compute_n_apply_grad1 (USE_GRAD1; the reference's GPU2
                             path), else compute_n_apply_grad2 (its CPU default) */
  int grad_calc_opt_mode; /* 2016 backend only: 0 DISABLED_OPT, 1 INTERNAL_OPT, 2 EXTERNAL_OPT
                             (factparams.py:952-962; pyfaust's default is 2) */
} fh_palm_params;

typedef struct fh_palm_result_s *fh_palm_result_t;

/* A is a host FP64 nrows x ncols col-major array. init_facts may be NULL. */
int fh_palm4msa(const double *A, int nrows, int ncols, const fh_constraint *cons, int ncons,
                const fh_stop_crit *sc, const fh_palm_params *p, double lambda_init, fh_faust_t init_facts,
                fh_palm_result_t *out);
int fh_hierarchical(const double *A, int nrows, int ncols, const fh_constraint *fac_cons,
                    const fh_constraint *res_cons, int nlevels, const fh_stop_crit *sc_local,
                    const fh_stop_crit *sc_global, const fh_palm_params *p, fh_palm_result_t *out);

/* The 2016 backend (pyfaust's default, backend=2016): Faust::Palm4MSA
 * (faust_Palm4MSA.hpp: dense factors, init_fact :628-665, next_step
 * :690-740, compute_grad_over_c :205-330, compute_lambda :333-358,
 * compute_c :612-625) and Faust::HierarchicalFact (faust_HierarchicalFact.hpp
 * :95-175).  Factors stay dense; the result's factors are the dense S_i and
 * its lambda is not folded in (FaustFact.hpp:380-420 and :486-500 build the
 * returned Faust from them).  Stopping: iteration counts only (an error
 * criterion gives FH_ERR_UNSUPPORTED).  The trajectory, when recorded, holds
 * (call, iteration, NaN, lambda) after every iteration. */
int fh_palm4msa2016(const double *A, int nrows, int ncols, const fh_constraint *cons, int ncons,
                    const fh_stop_crit *sc, const fh_palm_params *p, double lambda_init,
                    fh_faust_t init_facts, fh_palm_result_t *out);
int fh_hierarchical2016(const double *A, int nrows, int ncols, const fh_constraint *fac_cons,
                        const fh_constraint *res_cons, int nlevels, const fh_stop_crit *sc_local,
                        const fh_stop_crit *sc_global, const fh_palm_params *p, double init_lambda,
                        fh_palm_result_t *out);
int fh_palm_result_nfacts(fh_palm_result_t r, int *n);
int fh_palm_result_lambda(fh_palm_result_t r, double *lambda);
/* factor i as a new matrix handle (owned by the caller) */
int fh_palm_result_factor(fh_palm_result_t r, int i, fh_mat_t *out);
int fh_palm_result_traj(fh_palm_result_t r, double *out4, int *len);
int fh_palm_result_free(fh_palm_result_t r);

/* ---- butterfly factorization (Faust::butterfly_hierarchical,
 * faust_butterfly.hpp:804-860, as bound by FaustFact.hpp:905-926 and
 * _FaustAlgoGen.pyx.in:592-630).  Directions: ButterflyFactDir
 * (faust_butterfly.h:5-10); 'left' -> 0, 'right' -> 1, 'bbtree' -> 2. */
#define FH_BUTTERFLY_LEFT 0
#define FH_BUTTERFLY_RIGHT 1
#define FH_BUTTERFLY_BBTREE 2
/* A: n x n FP64, column-major, host memory, n = 2^L >= 4.  perm: NULL or n
 * indices of the permutation P with P[perm[i], i] = 1; then A P^T is
 * factorized and P is multiplied into the last factor (mul_perm != 0) or
 * appended as a factor of its own.  *out: a new Faust (owned by the caller)
 * of L factors on the butterfly supports (L + 1 with an appended
 * permutation): CSR factors for BBTREE, dense ones for LEFT / RIGHT (a
 * permuted last factor is CSR).  Every block is a best rank-1
 * approximation in FP64; BBTREE handles orders up to 32768 (the largest
 * block must fit one workgroup's LDS), LEFT / RIGHT any order. */
int fh_butterfly(const double *A, int n, int dir, const int32_t *perm, int mul_perm, fh_faust_t *out);
/* The same factorization of a matrix already in device memory: A column-major
 * (row_major = 0) or row-major (row_major = 1) with leading dimension lda;
 * perm (host memory) and the result as for fh_butterfly.  Work is enqueued
 * on `stream` and finished before the call returns.  This form has no
 * counterpart in the reference, whose butterfly_hierarchical takes host
 * matrices; it is for callers whose data is already on the GPU. */
int fh_butterfly_device(const double *A, int n, int64_t lda, int row_major, int dir, const int32_t *perm,
                        int mul_perm, fh_faust_t *out, void *stream);

/* ---- batched PALM4MSA: many independent factorizations in lock-step.
 * Every target b is n[b] x n[b] and is factorized into nfacts square factors
 * with palm4msa2's left-to-right, packed schedule (ParamsPalm4MSA defaults).
 * The main iterations run in `dtype` (FH_F32 or FH_F64); with refine_its > 0
 * and FH_F32 they are followed by FP64 sweeps in which each factor is
 * projected on its frozen support (prox supp, normalised).  Step sizes
 * (power iteration), lambda and the final relative errors are FP64.  One
 * iteration is captured once as a HIP graph and replayed when use_graph. */
typedef struct {
  int dtype;           /* FH_F32 or FH_F64 */
  int num_its;         /* main iterations */
  int refine_its;      /* FP64 frozen-support sweeps after FP32 iterations */
  int prox;            /* FH_PROX_SP (k per target) or FH_PROX_SUPP (masks) */
  int warm_power;      /* 1: start each power iteration from the previous vector */
  int norm2_max_iter;  /* default 100 */
  double norm2_threshold; /* default 1e-6 */
  int use_graph;       /* capture one iteration in a HIP graph */
  int autotune;        /* time 32/64 GEMM tiles per target size at set-up */
  int check_every;     /* > 0: FP64 relative-error check every check_every iterations */
  double target_err;   /* stop when every target's FP64 relative error <= target_err */
  const double *target_errs; /* per-target thresholds (ntargets of them) in place of target_err, or NULL */
} fh_batch_params;

typedef struct fh_batch_result_s *fh_batch_result_t;

/* A[b]: host FP64 n[b]*n[b] col-major; k[b]: sp parameter (FH_PROX_SP);
 * supp[b*nfacts+f]: host 0/1 masks (FH_PROX_SUPP); init[b*nfacts+f]: initial
 * factors or NULL for the palm4msa2 default (S_0 = 0, others identity);
 * lambda_init: per-target initial lambda or NULL (1). */
int fh_palm4msa_batched(int ntargets, int nfacts, const int *n, const double *const *A, const int64_t *k,
                        const double *const *supp, const double *const *init, const double *lambda_init,
                        const fh_batch_params *p, fh_batch_result_t *out);
int fh_batch_result_lambda(fh_batch_result_t r, double *lambdas);
/* FP64 relative error ||lambda S_0...S_{J-1} - A||_F / ||A||_F per target */
int fh_batch_result_relerr(fh_batch_result_t r, double *errs);
int fh_batch_result_factor(fh_batch_result_t r, int target, int f, double *dense_out);
/* ms[0] set-up, ms[1] main iterations, ms[2] refinement, ms[3] verification,
 * ms[4] wall time until every target reached its threshold (-1: not reached) */
int fh_batch_result_times(fh_batch_result_t r, double *ms5);
/* per target: iterations run when its error was first seen <= its threshold (-1: never) */
int fh_batch_result_iters(fh_batch_result_t r, int *iters);
int fh_batch_result_free(fh_batch_result_t r);

/* ---- warm-start MLP (n_in -> n_h1 -> n_h2 -> n_out, GELU between layers):
 * BF16 weights and activations on MFMA, FP32 accumulation, FP32 master
 * weights and gradients, Adam.  No counterpart in the reference: it proposes
 * initial factor values for PALM4MSA.  Matrices are row-major FP32 device
 * arrays (X: B x n_in, Y/T: B x n_out). */
typedef struct fh_mlp_s *fh_mlp_t;
int fh_mlp_create(int n_in, int n_h1, int n_h2, int n_out, uint64_t seed, fh_mlp_t *out);
int fh_mlp_free(fh_mlp_t M);
int fh_mlp_forward(fh_mlp_t M, const float *X, int B, float *Y, void *stream);
/* gradients of the mean squared error over one batch (loss to *loss; synchronizes) */
int fh_mlp_gradients(fh_mlp_t M, const float *X, const float *T, int B, float *loss, void *stream);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with learning rate lr */
int fh_mlp_train_step(fh_mlp_t M, const float *X, const float *T, int B, float lr, float *loss, void *stream);
/* host copies of layer l (0..2): what 0 = parameters, 1 = last gradients;
 * W is in x out row-major */
int fh_mlp_get(fh_mlp_t M, int layer, int what, float *W, float *b);
int fh_mlp_set(fh_mlp_t M, int layer, const float *W, const float *b);

#ifdef __cplusplus
}
#endif
#endif
