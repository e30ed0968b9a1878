// placeholder until the FP32/BF16 kernels land
#include "fh_api_common.hpp"
namespace fh {
void mixed_to_dense(const DMat&, void*, hipStream_t) { throw Error(FH_ERR_UNSUPPORTED, "FP32/BF16 not built"); }
void mixed_mul_dense(const DMat&, char, const void*, int, int, void*, int, const void*, const void*, hipStream_t) {
  throw Error(FH_ERR_UNSUPPORTED, "FP32/BF16 not built");
}
void faust_multiply_mixed(const std::vector<DMatP>&, char, const void*, int, void*, hipStream_t, DBuf*) {
  throw Error(FH_ERR_UNSUPPORTED, "FP32/BF16 not built");
}
}  // namespace fh
extern "C" int fh_mat_bsr_create(int, int, int, int, int, int, const int32_t*, const int32_t*, const void*, int,
                                 fh_mat_t*) {
  fh::set_last_error("BSR not built");
  return FH_ERR_UNSUPPORTED;
}
