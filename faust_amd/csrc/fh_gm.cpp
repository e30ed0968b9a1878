// fh_gm.cpp -- the gpu_mod function tables (include/faust_gm.h) over this
// library's handles and kernels: what the reference's GPU2 classes call
// through GPUModHandler (faust_MatDense_gpu.cpp.in, faust_MatSparse_gpu.cpp.in,
// faust_MatBSR_gpu.cpp.in), with their argument orders and ownership rules.
#include <cmath>

#include "../../include/faust_gm.h"
#include "fh_api_common.hpp"

using namespace fh;

namespace {

thread_local int g_gm_err = FH_OK;

// run f; on failure keep the message and the first status, return `fail`
template <class R, class F>
R guarded(R fail, F f) {
  try {
    return f();
  } catch (const Error& e) {
    set_last_error(e.what());
    if (g_gm_err == FH_OK) g_gm_err = e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    if (g_gm_err == FH_OK) g_gm_err = FH_ERR_NUMERIC;
  }
  return fail;
}
template <class F>
void guarded_void(F f) {
  guarded<int>(0, [&] {
    f();
    return 0;
  });
}

char op_char(int op) {
  if (op == FH_GM_OP_NOTRANSP) return 'N';
  if (op == FH_GM_OP_TRANSP) return 'T';
  if (op == FH_GM_OP_CONJTRANSP) return 'H';
  throw Error(FH_ERR_ARG, "invalid gm_Op value");
}

fh_mat_s* handle(fh_gm_mat_t p) {
  FH_CHECK(p, "null matrix handle");
  return static_cast<fh_mat_s*>(p);
}

template <class T> constexpr int dtype_of() { return std::is_same<T, double>::value ? FH_F64 : FH_F32; }

// roles the float tables do not serve
template <class T>
void require_f64(const char* what) {
  if (!std::is_same<T, double>::value) throw Error(FH_ERR_UNSUPPORTED, std::string(what) + ": FP64 only");
}

const DMat& dense_arg(fh_gm_mat_t p, int dtype) {
  const DMat& d = *handle(p)->m;
  FH_CHECK(d.kind == K_DENSE && d.dtype == dtype, "a dense handle of the table's scalar type is expected");
  return d;
}

DMatP new_dense(int dtype, int m, int n) {
  DMatP d = make_dense(dtype, m, n, nullptr, std::max(m, 1), false, nullptr);
  if ((size_t)m * n) FH_HIP(hipMemsetAsync(d->d.p, 0, dtype_size(dtype) * (size_t)m * n, nullptr));
  return d;
}

// B = A^T (column-major, ld = rows), 32 x 32 tiles through LDS
template <class T>
__global__ void k_gm_transpose(const T* __restrict__ A, int m, int n, T* __restrict__ B) {
  __shared__ T tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;  // bx: rows of A, by: columns of A
  for (int j = threadIdx.y; j < 32; j += 8) {
    const int r = bx + threadIdx.x, c = by + j;
    if (r < m && c < n) tile[j][threadIdx.x] = A[(size_t)c * m + r];
  }
  __syncthreads();
  for (int j = threadIdx.y; j < 32; j += 8) {
    const int r = by + threadIdx.x, c = bx + j;  // B is n x m
    if (r < n && c < m) B[(size_t)c * n + r] = tile[threadIdx.x][j];
  }
}

template <class T>
DMatP transposed(const DMat& A) {
  DMatP B = make_dense(A.dtype, A.n, A.m, nullptr, std::max(A.n, 1), false, nullptr);
  if ((size_t)A.m * A.n) {
    hipLaunchKernelGGL(k_gm_transpose<T>, dim3((A.m + 31) / 32, (A.n + 31) / 32), dim3(32, 8), 0, nullptr,
                       A.d.as<T>(), A.m, A.n, B->d.as<T>());
    FH_LAUNCH_CHECK();
  }
  return B;
}

// FP32 <-> BF16 (round to nearest even) for the BSR products of the float table
__global__ void k_gm_f32_to_bf16(const float* __restrict__ x, int64_t n, uint16_t* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = __builtin_bit_cast(uint16_t, (__bf16)x[i]);
}
__global__ void k_gm_bf16_to_f32(const uint16_t* __restrict__ x, int64_t n, float* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = __uint_as_float(((unsigned)x[i]) << 16);
}
void f32_to_bf16(const float* x, int64_t n, uint16_t* y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gm_f32_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, n, y);
  FH_LAUNCH_CHECK();
}
void bf16_to_f32(const uint16_t* x, int64_t n, float* y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gm_bf16_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, n, y);
  FH_LAUNCH_CHECK();
}

// the output of op(A) op(X): the given handle (resized when beta is 0 and
// its shape differs), or a new dense handle
fh_mat_s* output_of(fh_gm_mat_t out, int dtype, int rows, int cols, bool beta_zero) {
  if (!out) return new fh_mat_s{new_dense(dtype, rows, cols)};
  fh_mat_s* o = handle(out);
  FH_CHECK(o->m->kind == K_DENSE && o->m->dtype == dtype, "the output must be a dense handle of the table's type");
  if (o->m->m != rows || o->m->n != cols) {
    FH_CHECK(beta_zero, "output dimensions must agree when beta is not zero");
    o->m = new_dense(dtype, rows, cols);
  }
  return o;
}

// out = alpha op(S) op(X) + beta out, S any matrix handle of the scalar type
template <class T>
fh_gm_mat_t mul_ext(fh_gm_mat_t S, fh_gm_mat_t X, fh_gm_mat_t out, int opS, int opX, const T* alpha, const T* beta) {
  const DMat& Sm = *handle(S)->m;
  const DMat& Xm = dense_arg(X, dtype_of<T>());
  const char os = op_char(opS), ox = op_char(opX);
  const int rows = os == 'N' ? Sm.m : Sm.n, inner = os == 'N' ? Sm.n : Sm.m;
  const int xr = ox == 'N' ? Xm.m : Xm.n, cols = ox == 'N' ? Xm.n : Xm.m;
  FH_CHECK(xr == inner, "mul_gpu_dsm_ext: dimensions do not agree");
  const T a = alpha ? *alpha : T(1), b = beta ? *beta : T(0);
  fh_mat_s* o = output_of(out, dtype_of<T>(), rows, cols, b == T(0));
  DMatP xt;
  const DMat* xin = &Xm;
  if (ox != 'N') {
    xt = transposed<T>(Xm);
    xin = xt.get();
  }
  if (cols > 0 && rows > 0) {
    if (Sm.dtype == FH_F64) {
      FH_CHECK((std::is_same<T, double>::value), "an FP64 matrix in a float table");
      if (Sm.kind == K_CSR)
        csr_spmm_f64(Sm, os, xin->d.as<double>(), std::max(inner, 1), cols, o->m->d.as<double>(), rows, (double)a,
                     (double)b, nullptr);
      else
        gemm_f64(os == 'N' ? 'N' : 'T', 'N', rows, cols, inner, (double)a, Sm.d.as<double>(), Sm.m,
                 xin->d.as<double>(), std::max(inner, 1), (double)b, o->m->d.as<double>(), rows, nullptr);
    } else if (Sm.kind == K_BSR && Sm.dtype == FH_F32) {
      // FP32 blocks, FP32 activations (MatBSR<float>::multiply's precision)
      bsr_mul_f32(Sm, os, xin->d.as<float>(), std::max(inner, 1), cols, o->m->d.as<float>(), rows, (float)a,
                  (float)b, nullptr);
    } else if (Sm.kind == K_BSR) {
      // BF16 blocks: BF16 activations on MFMA, converted around the product
      DBuf xb(sizeof(uint16_t) * std::max<size_t>((size_t)inner * cols, 1)),
          yb(sizeof(uint16_t) * std::max<size_t>((size_t)rows * cols, 1));
      f32_to_bf16(xin->d.as<float>(), (int64_t)inner * cols, xb.as<uint16_t>(), nullptr);
      if (b != T(0)) f32_to_bf16(o->m->d.as<float>(), (int64_t)rows * cols, yb.as<uint16_t>(), nullptr);
      const float fa = (float)a, fb = (float)b;
      mixed_mul_dense(Sm, os, xb.p, inner, cols, yb.p, rows, &fa, &fb, nullptr);
      bf16_to_f32(yb.as<uint16_t>(), (int64_t)rows * cols, o->m->d.as<float>(), nullptr);
      FH_HIP(hipStreamSynchronize(nullptr));  // xb / yb go back to the pool
    } else {
      const float fa = (float)a, fb = (float)b;
      mixed_mul_dense(Sm, os, xin->d.p, std::max(inner, 1), cols, o->m->d.p, rows, &fa, &fb, nullptr);
    }
  }
  if (xt) FH_HIP(hipStreamSynchronize(nullptr));  // the transposed copy goes back to the pool
  return o;
}

// ---- dense
template <class T>
fh_gm_mat_t dsm_create(int32_t m, int32_t n) {
  return guarded<fh_gm_mat_t>(nullptr, [&] { return (fh_gm_mat_t) new fh_mat_s{new_dense(dtype_of<T>(), m, n)}; });
}
template <class T>
fh_gm_mat_t dsm_togpu(int32_t m, int32_t n, const T* data) {
  return guarded<fh_gm_mat_t>(nullptr, [&] {
    FH_CHECK(m >= 0 && n >= 0 && (data || (size_t)m * n == 0), "togpu: bad arguments");
    return (fh_gm_mat_t) new fh_mat_s{make_dense(dtype_of<T>(), m, n, data, std::max(m, 1), false, nullptr)};
  });
}
template <class T>
void dsm_tocpu(fh_gm_mat_t M, T* data) {
  guarded_void([&] {
    const DMat& d = dense_arg(M, dtype_of<T>());
    FH_HIP(hipStreamSynchronize(nullptr));
    if ((size_t)d.m * d.n) FH_HIP(hipMemcpy(data, d.d.p, sizeof(T) * (size_t)d.m * d.n, hipMemcpyDeviceToHost));
  });
}
void mat_free(fh_gm_mat_t M) {
  guarded_void([&] {
    FH_HIP(hipStreamSynchronize(nullptr));  // kernels still reading the matrix finish first
    delete static_cast<fh_mat_s*>(M);
  });
}
void dsm_info(fh_gm_mat_t M, int32_t* m, int32_t* n) {
  guarded_void([&] {
    const DMat& d = *handle(M)->m;
    if (m) *m = d.m;
    if (n) *n = d.n;
  });
}
template <class T>
void dsm_gemm(fh_gm_mat_t A, fh_gm_mat_t B, fh_gm_mat_t C, const T* alpha, const T* beta, int opA, int opB) {
  guarded_void([&] {
    require_f64<T>("gemm");
    FH_CHECK(C, "gemm: null output");
    const DMat& a = dense_arg(A, FH_F64);
    const DMat& b = dense_arg(B, FH_F64);
    const char oa = op_char(opA), ob = op_char(opB);
    const int m = oa == 'N' ? a.m : a.n, k = oa == 'N' ? a.n : a.m;
    const int kb = ob == 'N' ? b.m : b.n, n = ob == 'N' ? b.n : b.m;
    FH_CHECK(k == kb, "gemm: dimensions do not agree");
    const double al = alpha ? (double)*alpha : 1.0, be = beta ? (double)*beta : 0.0;
    fh_mat_s* c = output_of(C, FH_F64, m, n, be == 0.0);
    if (m > 0 && n > 0)
      gemm_f64(oa == 'N' ? 'N' : 'T', ob == 'N' ? 'N' : 'T', m, n, k, al, a.d.as<double>(), std::max(a.m, 1),
               b.d.as<double>(), std::max(b.m, 1), be, c->m->d.as<double>(), std::max(m, 1), nullptr);
  });
}
template <class T>
fh_gm_mat_t dsm_mul_ext(fh_gm_mat_t A, fh_gm_mat_t B, fh_gm_mat_t out, int opA, int opB) {
  return guarded<fh_gm_mat_t>(nullptr, [&] {
    require_f64<T>("dense products");
    dense_arg(A, FH_F64);
    return mul_ext<T>(A, B, out, opA, opB, nullptr, nullptr);
  });
}
template <class T, int NAME>
void dsm_prox(fh_gm_mat_t M, int32_t k, bool normalized, bool pos) {
  guarded_void([&] {
    require_f64<T>("projectors");
    const DMat& d = dense_arg(M, FH_F64);
    prox_f64(NAME, d.d.as<double>(), d.m, d.n, k, normalized, pos, nullptr, nullptr);
  });
}
template <class T>
T dsm_norm_spectral(fh_gm_mat_t M, float threshold, int32_t max_iter) {
  return guarded<T>(T(0), [&] {
    require_f64<T>("norm_spectral");
    dense_arg(M, FH_F64);
    int flag = 0;
    return (T)faust_spectral_norm_f64({handle(M)->m}, max_iter, threshold, &flag, nullptr);
  });
}
template <class T>
T dsm_norm_frob(fh_gm_mat_t M) {
  return guarded<T>(T(0), [&] {
    require_f64<T>("norm_frob");
    const DMat& d = dense_arg(M, FH_F64);
    DBuf s(sizeof(double));
    reduce_dot_f64(d.d.as<double>(), d.d.as<double>(), (int64_t)d.m * d.n, s.as<double>(), nullptr);
    double h = 0;
    FH_HIP(hipMemcpy(&h, s.p, sizeof(double), hipMemcpyDeviceToHost));
    return (T)std::sqrt(h);
  });
}

// ---- sparse (CSR)
template <class T>
fh_gm_mat_t spm_togpu(int32_t m, int32_t n, int32_t nnz, const int32_t* rp, const int32_t* ci, const T* v) {
  return guarded<fh_gm_mat_t>(nullptr, [&] {
    FH_CHECK(rp && (nnz == 0 || (ci && v)), "togpu: null arrays");
    return (fh_gm_mat_t) new fh_mat_s{make_csr_from_host(dtype_of<T>(), m, n, nnz, rp, ci, v, nullptr)};
  });
}
template <class T>
void spm_tocpu(fh_gm_mat_t S, int32_t* rp, int32_t* ci, T* v) {
  guarded_void([&] {
    const DMat& d = *handle(S)->m;
    FH_CHECK(d.kind == K_CSR && d.dtype == dtype_of<T>(), "tocpu: a CSR handle of the table's type is expected");
    FH_HIP(hipStreamSynchronize(nullptr));
    if (rp) FH_HIP(hipMemcpy(rp, d.rp.p, sizeof(int) * (d.m + 1), hipMemcpyDeviceToHost));
    if (d.nnz) {
      if (ci) FH_HIP(hipMemcpy(ci, d.ci.p, sizeof(int) * d.nnz, hipMemcpyDeviceToHost));
      if (v) FH_HIP(hipMemcpy(v, d.v.p, sizeof(T) * d.nnz, hipMemcpyDeviceToHost));
    }
  });
}
void spm_info(fh_gm_mat_t S, int32_t* m, int32_t* n, int32_t* nnz) {
  guarded_void([&] {
    const DMat& d = *handle(S)->m;
    if (m) *m = d.m;
    if (n) *n = d.n;
    if (nnz) *nnz = (int32_t)d.nnz;
  });
}
template <class T>
fh_gm_mat_t spm_mul_ext(fh_gm_mat_t S, fh_gm_mat_t X, fh_gm_mat_t out, int opS, int opX) {
  return guarded<fh_gm_mat_t>(nullptr, [&] {
    FH_CHECK(handle(S)->m->kind == K_CSR, "a sparse handle is expected");
    return mul_ext<T>(S, X, out, opS, opX, nullptr, nullptr);
  });
}
template <class T>
void spm_gemm(fh_gm_mat_t A, fh_gm_mat_t B, fh_gm_mat_t C, const T* alpha, const T* beta, int opA, int opB) {
  guarded_void([&] {
    FH_CHECK(C && handle(A)->m->kind == K_CSR, "spgemm: a sparse A and an output are expected");
    mul_ext<T>(A, B, C, opA, opB, alpha, beta);
  });
}

// ---- BSR
template <class T>
fh_gm_mat_t bsr_togpu(int32_t m, int32_t n, int32_t bm, int32_t bn, const T* bdata, const int32_t* browptr,
                      const int32_t* bcolinds, int32_t bnnz) {
  return guarded<fh_gm_mat_t>(nullptr, [&] {
    fh_mat_t h = nullptr;
    const int rc = fh_mat_bsr_create(dtype_of<T>(), m, n, bm, bn, bnnz, browptr, bcolinds, bdata, 0, &h);
    if (rc != FH_OK) throw Error(rc, fh_last_error());
    return (fh_gm_mat_t)h;
  });
}
void bsr_info(fh_gm_mat_t B, int32_t* m, int32_t* n, int32_t* bm, int32_t* bn, int32_t* bnnz) {
  guarded_void([&] {
    const DMat& d = *handle(B)->m;
    if (m) *m = d.m;
    if (n) *n = d.n;
    if (d.kind == K_BSR) {
      if (bm) *bm = 32;
      if (bn) *bn = 32;
      if (bnnz) *bnnz = d.bnnz;
    } else {  // an FP64 BSR factor kept as CSR: its nonzeros
      if (bm) *bm = 1;
      if (bn) *bn = 1;
      if (bnnz) *bnnz = (int32_t)d.nnz;
    }
  });
}
template <class T>
fh_gm_mat_t bsr_mul_ext(fh_gm_mat_t B, fh_gm_mat_t X, fh_gm_mat_t out, int opB, int opX, const T* alpha,
                        const T* beta) {
  return guarded<fh_gm_mat_t>(nullptr, [&] { return mul_ext<T>(B, X, out, opB, opX, alpha, beta); });
}

int32_t gp_dev_count() {
  return guarded<int32_t>(0, [] {
    int c = 0;
    FH_HIP(hipGetDeviceCount(&c));
    return (int32_t)c;
  });
}
void gp_set_dev(int32_t d) { guarded_void([&] { FH_HIP(hipSetDevice(d)); }); }
int32_t gp_cur_dev() {
  return guarded<int32_t>(-1, [] {
    int d = 0;
    FH_HIP(hipGetDevice(&d));
    return (int32_t)d;
  });
}

template <class T, class Table>
void fill_dense(Table* t) {
  t->create = dsm_create<T>;
  t->togpu = dsm_togpu<T>;
  t->tocpu = dsm_tocpu<T>;
  t->free = mat_free;
  t->info = dsm_info;
  t->gemm = dsm_gemm<T>;
  t->mul_gpu_dsm_ext = dsm_mul_ext<T>;
  t->prox_sp = dsm_prox<T, FH_PROX_SP>;
  t->prox_spcol = dsm_prox<T, FH_PROX_SPCOL>;
  t->prox_splin = dsm_prox<T, FH_PROX_SPLIN>;
  t->prox_splincol = dsm_prox<T, FH_PROX_SPLINCOL>;
  t->norm_spectral = dsm_norm_spectral<T>;
  t->norm_frob = dsm_norm_frob<T>;
}
template <class T, class Table>
void fill_sparse(Table* t) {
  t->togpu = spm_togpu<T>;
  t->tocpu = spm_tocpu<T>;
  t->free = mat_free;
  t->info = spm_info;
  t->mul_gpu_dsm_ext = spm_mul_ext<T>;
  t->gemm = spm_gemm<T>;
}
template <class T, class Table>
void fill_bsr(Table* t) {
  t->togpu = bsr_togpu<T>;
  t->free = mat_free;
  t->info = bsr_info;
  t->mul_gpu_dsm_ext = bsr_mul_ext<T>;
}

}  // namespace

extern "C" {
void fh_gm_load_dsm_funcs_double(fh_gm_DenseMatFunc_double* t) { fill_dense<double>(t); }
void fh_gm_load_dsm_funcs_float(fh_gm_DenseMatFunc_float* t) { fill_dense<float>(t); }
void fh_gm_load_spm_funcs_double(fh_gm_SparseMatFunc_double* t) { fill_sparse<double>(t); }
void fh_gm_load_spm_funcs_float(fh_gm_SparseMatFunc_float* t) { fill_sparse<float>(t); }
void fh_gm_load_bsr_funcs_double(fh_gm_BSRMatFunc_double* t) { fill_bsr<double>(t); }
void fh_gm_load_bsr_funcs_float(fh_gm_BSRMatFunc_float* t) { fill_bsr<float>(t); }
void fh_gm_load_gp_funcs(fh_gm_GenPurposeFunc* t) {
  t->dev_count = gp_dev_count;
  t->set_dev = gp_set_dev;
  t->cur_dev = gp_cur_dev;
}
int fh_gm_take_error(void) {
  const int e = g_gm_err;
  g_gm_err = FH_OK;
  return e;
}
}
