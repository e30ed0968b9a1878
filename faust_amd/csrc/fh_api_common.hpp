// fh_api_common.hpp -- definitions shared by the C ABI translation units.
#pragma once
#include <cmath>

#include "fh_internal.hpp"

struct fh_mat_s {
  fh::DMatP m;
};
// A Faust keeps its two ping-pong intermediates (Transform::multiply's
// tmp_buf1/tmp_buf2, faust_Transform.hpp:1780-1782) between calls, so a
// repeated product allocates nothing, synchronizes nothing and can be
// captured in a HIP graph.
struct fh_faust_s {
  std::vector<fh::DMatP> f;
  fh::DBuf ws[2];
};

#define FH_API_BEGIN try {
#define FH_API_END                 \
  return FH_OK;                    \
  }                                \
  catch (const fh::Error& e) {     \
    fh::set_last_error(e.what());  \
    return e.code;                 \
  }                                \
  catch (const std::exception& e) { \
    fh::set_last_error(e.what());  \
    return FH_ERR_NUMERIC;         \
  }

namespace fh {
DMatP make_csr_from_host(int dtype, int m, int n, int nnz, const int32_t* rp, const int32_t* ci, const void* vals,
                         hipStream_t st);
DMatP make_dense(int dtype, int m, int n, const void* data, int ld, bool on_device, hipStream_t st);
void faust_multiply_f64(const std::vector<DMatP>& f, char op, const double* X, int p, double* Y, hipStream_t st,
                        DBuf* ws);
DMatP faust_get_product_f64(const std::vector<DMatP>& f, hipStream_t st);
double faust_spectral_norm_f64(const std::vector<DMatP>& f, int maxit, double thr, int* flag, hipStream_t st);
// FP32 / BF16 paths (k_mixed.hip, k_bsr.hip)
void mixed_to_dense(const DMat& M, void* dense_dev, hipStream_t st);
void mixed_mul_dense(const DMat& S, char op, const void* X, int ldx, int ncols, void* Y, int ldy, const void* alpha,
                     const void* beta, hipStream_t st);
void faust_multiply_mixed(const std::vector<DMatP>& f, char op, const void* X, int p, void* Y, hipStream_t st,
                          DBuf* ws);
}  // namespace fh
