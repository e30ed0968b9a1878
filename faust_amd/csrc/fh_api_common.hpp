// fh_api_common.hpp -- definitions shared by the C ABI translation units.
#pragma once
#include <cmath>

#include "fh_internal.hpp"

struct fh_mat_s {
  fh::DMatP m;
};
// Per-factor device timers (the reference's per-factor Transform::multiply
// timers, faust_Transform.hpp:1499-1536): HIP events recorded on the launch
// stream around each factor's kernel; slot nfactors collects layout
// conversions.
namespace fh {
struct FaustTimers {
  bool on = false;
  std::vector<std::vector<std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<double> ms;
  std::vector<int64_t> count;
  std::vector<hipEvent_t> spare;
  void resize(size_t slots) {
    if (pending.size() < slots) {
      pending.resize(slots);
      ms.resize(slots, 0.0);
      count.resize(slots, 0);
    }
  }
  hipEvent_t get() {
    if (!spare.empty()) {
      hipEvent_t e = spare.back();
      spare.pop_back();
      return e;
    }
    hipEvent_t e;
    FH_HIP(hipEventCreate(&e));
    return e;
  }
  void begin(size_t slot, hipStream_t st) {
    if (!on) return;
    resize(slot + 1);
    hipEvent_t a = get();
    FH_HIP(hipEventRecord(a, st));
    pending[slot].push_back({a, nullptr});
  }
  void end(size_t slot, hipStream_t st) {
    if (!on) return;
    hipEvent_t b = get();
    FH_HIP(hipEventRecord(b, st));
    pending[slot].back().second = b;
  }
  void collect() {
    for (size_t s = 0; s < pending.size(); ++s) {
      for (auto& pr : pending[s]) {
        FH_HIP(hipEventSynchronize(pr.second));
        float t = 0.f;
        FH_HIP(hipEventElapsedTime(&t, pr.first, pr.second));
        ms[s] += t;
        count[s] += 1;
        spare.push_back(pr.first);
        spare.push_back(pr.second);
      }
      pending[s].clear();
    }
  }
  ~FaustTimers() {
    for (auto& v : pending)
      for (auto& pr : v) {
        (void)hipEventDestroy(pr.first);
        if (pr.second) (void)hipEventDestroy(pr.second);
      }
    for (auto e : spare) (void)hipEventDestroy(e);
  }
};
}  // namespace fh

// A Faust keeps its two ping-pong intermediates (Transform::multiply's
// tmp_buf1/tmp_buf2, faust_Transform.hpp:1780-1782) between calls, so a
// repeated product allocates nothing, synchronizes nothing and can be
// captured in a HIP graph.
struct fh_faust_s {
  std::vector<fh::DMatP> f;
  fh::DBuf ws[2];         // ping-pong activations
  fh::DBuf ctl;           // control block of the one-launch BSR chain
  int last_launches = 0;  // kernel launches of the last product
  int mul_mode = 0;       // FH_MUL_*
  int act_dtype = -1;     // activations of a BSR chain: FH_BF16 (default) or FH_F32 (FP32 BSR factors)
  fh::FaustTimers timers;
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
// The intermediates of one chain product: device addresses, each at least the
// bytes faust_ws_need_* gives (unused entries may be null).
struct ChainWs {
  void* p[2];
  int* ctl;  // control block of the one-launch BSR chain (bsr_chain_ctl_bytes; zeroed when allocated), or null
};
size_t faust_ws_need_f64(const std::vector<DMatP>& f, char op, int p, int* nbuf);
ChainWs faust_ws_grow(DBuf* ws, size_t need, int nbuf, hipStream_t st, DBuf* ctl = nullptr);
// Forgets the forms measured by the chain autotuners (FP64 and BSR / FP32 CSR).
void autotune_clear();
// Shared by the autotuners: FH_AUTOTUNE=0 turns measuring off; a stream being
// captured into a graph is never measured on; the BSR / FP32 CSR store maps a
// chain signature to the FH_MUL_* mode measured quickest.
bool autotune_on();
bool stream_capturing(hipStream_t st);
bool autotune_get_mode(const std::string& key, int* mode);
void autotune_put_mode(const std::string& key, int mode);
void faust_multiply_f64(const std::vector<DMatP>& f, char op, const double* X, int p, double* Y, hipStream_t st,
                        const ChainWs& ws, FaustTimers* tm = nullptr, int* launches = nullptr, bool tune = true);
DMatP faust_get_product_f64(const std::vector<DMatP>& f, hipStream_t st);
double faust_spectral_norm_f64(const std::vector<DMatP>& f, int maxit, double thr, int* flag, hipStream_t st);
void faust_spectral_norm_pair_f64(const std::vector<DMatP>& f1, const std::vector<DMatP>& f2, int maxit, double thr,
                                  int* flag, double* n1, double* n2, hipStream_t st);
// FP32 / BF16 paths (k_mixed.hip, k_bsr.hip)
void mixed_to_dense(const DMat& M, void* dense_dev, hipStream_t st);
void bsr_to_host(const DMat& M, void* bdata, int32_t* browptr, int32_t* bcolinds);
void mixed_mul_dense(const DMat& S, char op, const void* X, int ldx, int ncols, void* Y, int ldy, const void* alpha,
                     const void* beta, hipStream_t st);
size_t faust_ws_need_mixed(const std::vector<DMatP>& f, char op, int p, int* nbuf, bool act_f32 = false);
// one-launch BSR chain (k_bsr_chain.hip)
size_t bsr_chain_ctl_bytes();
void bsr_chain_ctl_clear(int* ctl, hipStream_t st);
int bsr_chain_report(int* ctl, hipStream_t st);
int bsr_chain_sync_mode();
bool bsr_chain_fits(const std::vector<DMatP>& f, char op);
void bsr_chain_launch(const std::vector<DMatP>& f, char op, const void* X, int p, void* Y, void* T0, void* T1,
                      int* ctl, hipStream_t st);
int faust_multiply_mixed(const std::vector<DMatP>& f, char op, const void* X, int p, void* Y, hipStream_t st,
                          const ChainWs& ws, FaustTimers* tm = nullptr, bool act_f32 = false, bool tune = true);
// Y = alpha op(B) X + beta Y, B an FP32 BSR factor, X / Y FP32 column-major
void bsr_mul_f32(const DMat& B, char op, const float* X, int ldx, int p, float* Y, int ldy, float alpha, float beta,
                 hipStream_t st);
}  // namespace fh
