/*
 * faust_gm.h -- the gpu_mod function tables of the reference's GPU2 layer,
 * served by libfaust_hip.so.
 *
 * The reference's Faust::GPUModHandler (src/faust_linear_operator/GPU2/
 * faust_gpu_mod_utils.h:24-97, .cpp:40-90) dlopens libgm.so and fills one
 * function table per role and scalar type (gm_DenseMatFunc_double,
 * gm_SparseMatFunc_double, gm_BSRMatFunc_double, ... and gm_GenPurposeFunc)
 * with load_dsm_funcs_double(handle, table) and its siblings.  The tables
 * below have the members the GPU2 classes call, with the argument order of
 * those call sites (faust_MatDense_gpu.cpp.in, faust_MatSparse_gpu.cpp.in,
 * faust_MatBSR_gpu.cpp.in), and fh_gm_load_*_funcs_* fill them.
 *
 * Conventions kept from the call sites:
 *   - handles are opaque pointers owned by the caller's matrix objects; a
 *     handle of this library is an fh_mat_t (faust_hip.h);
 *   - dense matrices are column-major with ld = nrows;
 *   - mul_gpu_dsm_ext(A, X, out, opA, opX[, alpha, beta]) returns out, or a
 *     new dense handle when out is NULL (the caller then frees the input,
 *     faust_MatSparse_gpu.cpp.in:462-472); a given out whose shape differs
 *     is resized when beta is 0 (its contents are not read);
 *   - alpha / beta are passed by pointer, NULL meaning 1 / 0
 *     (faust_MatBSR_gpu.cpp.in:115);
 *   - op values: FH_GM_OP_NOTRANSP / _TRANSP / _CONJTRANSP for 'N', 'T', 'H'
 *     (char2gm_Op, faust_gpu_mod_utils.cpp:189-199); the scalars are real,
 *     so 'H' acts as 'T'.
 * gpu_mod's functions return pointers or nothing; a failed call here returns
 * NULL (or 0 for scalars, or does nothing), keeps the message for
 * fh_last_error() and records its FH_ERR_* status for fh_gm_take_error().
 * All work is ordered on the null stream; tocpu waits for it.
 *
 * Coverage: double tables hold every role below; the float tables hold
 * handles and products (FP32 CSR; FP32 BSR blocks on FP32 MFMA with FP32
 * activations), and their gemm / prox / norm entries fail with
 * FH_ERR_UNSUPPORTED.
 */
#ifndef FAUST_GM_H
#define FAUST_GM_H
#include <stdbool.h>
#include <stdint.h>

#include "faust_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { FH_GM_OP_NOTRANSP = 0, FH_GM_OP_TRANSP = 1, FH_GM_OP_CONJTRANSP = 2 };
typedef void *fh_gm_mat_t; /* gm_DenseMat_t / gm_SparseMat_t / gm_BSRMat_t */

#define FH_GM_DENSE_TABLE(T)                                                                          \
  typedef struct {                                                                                    \
    fh_gm_mat_t (*create)(int32_t nrows, int32_t ncols);                                              \
    fh_gm_mat_t (*togpu)(int32_t nrows, int32_t ncols, const T *data);                                \
    void (*tocpu)(fh_gm_mat_t M, T *data);                                                            \
    void (*free)(fh_gm_mat_t M);                                                                      \
    void (*info)(fh_gm_mat_t M, int32_t *nrows, int32_t *ncols);                                      \
    void (*gemm)(fh_gm_mat_t A, fh_gm_mat_t B, fh_gm_mat_t C, const T *alpha, const T *beta, int opA,  \
                 int opB);                                                                            \
    fh_gm_mat_t (*mul_gpu_dsm_ext)(fh_gm_mat_t A, fh_gm_mat_t B, fh_gm_mat_t out, int opA, int opB);  \
    void (*prox_sp)(fh_gm_mat_t M, int32_t k, bool normalized, bool pos);                             \
    void (*prox_spcol)(fh_gm_mat_t M, int32_t k, bool normalized, bool pos);                          \
    void (*prox_splin)(fh_gm_mat_t M, int32_t k, bool normalized, bool pos);                          \
    void (*prox_splincol)(fh_gm_mat_t M, int32_t k, bool normalized, bool pos);                       \
    T (*norm_spectral)(fh_gm_mat_t M, float threshold, int32_t max_iter);                             \
    T (*norm_frob)(fh_gm_mat_t M);                                                                    \
  } fh_gm_DenseMatFunc_##T;

#define FH_GM_SPARSE_TABLE(T)                                                                         \
  typedef struct {                                                                                    \
    fh_gm_mat_t (*togpu)(int32_t nrows, int32_t ncols, int32_t nnz, const int32_t *rowptr,            \
                         const int32_t *colind, const T *values);                                     \
    void (*tocpu)(fh_gm_mat_t S, int32_t *rowptr, int32_t *colind, T *values);                        \
    void (*free)(fh_gm_mat_t S);                                                                      \
    void (*info)(fh_gm_mat_t S, int32_t *nrows, int32_t *ncols, int32_t *nnz);                        \
    fh_gm_mat_t (*mul_gpu_dsm_ext)(fh_gm_mat_t S, fh_gm_mat_t X, fh_gm_mat_t out, int opS, int opX);  \
    void (*gemm)(fh_gm_mat_t A, fh_gm_mat_t B, fh_gm_mat_t C, const T *alpha, const T *beta, int opA,  \
                 int opB);                                                                            \
  } fh_gm_SparseMatFunc_##T;

#define FH_GM_BSR_TABLE(T)                                                                            \
  typedef struct {                                                                                    \
    fh_gm_mat_t (*togpu)(int32_t nrows, int32_t ncols, int32_t bnrows, int32_t bncols, const T *bdata, \
                         const int32_t *browptr, const int32_t *bcolinds, int32_t bnnz);              \
    void (*free)(fh_gm_mat_t B);                                                                      \
    void (*info)(fh_gm_mat_t B, int32_t *nrows, int32_t *ncols, int32_t *bnrows, int32_t *bncols,     \
                 int32_t *bnnz);                                                                      \
    fh_gm_mat_t (*mul_gpu_dsm_ext)(fh_gm_mat_t B, fh_gm_mat_t X, fh_gm_mat_t out, int opB, int opX,    \
                                   const T *alpha, const T *beta);                                    \
  } fh_gm_BSRMatFunc_##T;

FH_GM_DENSE_TABLE(double)
FH_GM_DENSE_TABLE(float)
FH_GM_SPARSE_TABLE(double)
FH_GM_SPARSE_TABLE(float)
FH_GM_BSR_TABLE(double)
FH_GM_BSR_TABLE(float)

typedef struct {
  int32_t (*dev_count)(void);
  void (*set_dev)(int32_t dev);
  int32_t (*cur_dev)(void);
} fh_gm_GenPurposeFunc;

/* load_dsm_funcs_double / load_spm_funcs_double / load_bsr_funcs_double /
 * load_gp_funcs and their float siblings (faust_gpu_mod_utils.cpp:40-90). */
void fh_gm_load_dsm_funcs_double(fh_gm_DenseMatFunc_double *t);
void fh_gm_load_dsm_funcs_float(fh_gm_DenseMatFunc_float *t);
void fh_gm_load_spm_funcs_double(fh_gm_SparseMatFunc_double *t);
void fh_gm_load_spm_funcs_float(fh_gm_SparseMatFunc_float *t);
void fh_gm_load_bsr_funcs_double(fh_gm_BSRMatFunc_double *t);
void fh_gm_load_bsr_funcs_float(fh_gm_BSRMatFunc_float *t);
void fh_gm_load_gp_funcs(fh_gm_GenPurposeFunc *t);

/* Status of the first table call that failed on this thread since the last
 * take (FH_OK if none), then clears it; the message is fh_last_error(). */
int fh_gm_take_error(void);

#ifdef __cplusplus
}
#endif
#endif
