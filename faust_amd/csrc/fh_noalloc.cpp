// fh_noalloc.cpp -- fh_faust_multiply_noalloc: the Faust chain product with
// its two ping-pong intermediates in a fixed device array, for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  The kernels are those of
// fh_faust_multiply; only where the intermediates live differs, so both
// entry points give the same bits.
//
// `scratch` is one FH_NOALLOC_SCRATCH_BYTES array per GPU, held from module
// load to process end; BSR chains keep the control block of the one-launch
// chain at its start.  In host code its name is only a host-side stand-in
// for the symbol: the device address comes from hipGetSymbolAddress, once per
// device (with that device current), and only that address is handed to
// kernels.  Every call on one GPU shares the array; calls that overlap on one
// GPU (two streams, two threads) race on it.  The intermediates are written
// before they are read within a call, and the chain's counters are cleared by
// the chain itself at its end, so nothing needs clearing between calls.
#include <mutex>

#include "fh_api_common.hpp"

namespace {

constexpr size_t kScratchBytes = FH_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of the current device's scratch array, looked up on the
// first call on that device.
void* scratch_address() {
  int dev = 0;
  FH_HIP(hipGetDevice(&dev));
  FH_CHECK(dev >= 0 && dev < kMaxDevices, "fh_faust_multiply_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
      throw fh::Error(FH_ERR_HIP, std::string("fh_faust_multiply_noalloc: hipGetSymbolAddress(scratch): ") +
                                      hipGetErrorString(e));
    g_scratch[dev] = ptr;
  }
  return g_scratch[dev];
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

using namespace fh;

extern "C" int fh_faust_multiply_noalloc(fh_faust_t F, char op, const void* X, int p, void* Y, void* stream) {
  FH_API_BEGIN
  FH_CHECK(F, "null argument");
  MulModeScope mode_scope(F->mul_mode);
  FH_CHECK(op == 'N' || op == 'T' || op == 'H', "op must be N, T or H");
  FH_CHECK(!F->f.empty(), "Faust::multiply: empty Faust");
  FH_CHECK(p >= 0, "negative column count");
  if (p == 0) {
    F->last_launches = 0;
    return FH_OK;
  }
  FH_CHECK(X && Y, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const bool f64 = F->f.front()->dtype == FH_F64;
  // FP32 chains run dense FP32 factors as their CSR twins; building a twin
  // allocates, so it must already exist (fh_faust_multiply builds it)
  std::vector<DMatP> f = F->f;
  if (!f64)
    for (auto& x : f)
      if (x->kind == K_DENSE && x->dtype == FH_F32) {
        if (!x->csr_twin)
          throw Error(FH_ERR_UNSUPPORTED,
                      "fh_faust_multiply_noalloc: a dense FP32 factor has no CSR twin yet (one fh_faust_multiply "
                      "call on this Faust builds it)");
        x = x->csr_twin;
      }
  int nbuf = 0;
  const bool act_f32 = F->act_dtype == FH_F32;
  const size_t need = f64 ? faust_ws_need_f64(f, op, p, &nbuf) : faust_ws_need_mixed(f, op, p, &nbuf, act_f32);
  const size_t slot = align256(need);
  // BSR chains also keep the one-launch chain's control block at the start of
  // the array (zero at load; every chain leaves its counters cleared)
  const bool bsr = f.front()->kind == K_BSR && !act_f32;
  const size_t ctl_bytes = bsr ? align256(bsr_chain_ctl_bytes()) : 0;
  if (ctl_bytes + (size_t)nbuf * slot > kScratchBytes)
    throw Error(FH_ERR_ARG, "fh_faust_multiply_noalloc: this product needs " + std::to_string(nbuf) + " x " +
                                std::to_string(need) + " bytes of intermediates, more than the " +
                                std::to_string(kScratchBytes) + "-byte scratch array");
  char* base = (nbuf > 0 || bsr) ? static_cast<char*>(scratch_address()) : nullptr;
  const ChainWs ws{{nbuf > 0 ? base + ctl_bytes : nullptr, nbuf > 1 ? base + ctl_bytes + slot : nullptr},
                   bsr ? reinterpret_cast<int*>(base) : nullptr};
  if (f64)
    faust_multiply_f64(f, op, (const double*)X, p, (double*)Y, st, ws, nullptr, &F->last_launches, false);
  else
    F->last_launches = faust_multiply_mixed(f, op, X, p, Y, st, ws, nullptr, act_f32, false);
  FH_API_END
}
