// k_bsr_chain.hip -- a whole BSR Faust x dense chain (Transform::multiply,
// faust_Transform.hpp:1714-1812) in one launch.
//
// Column ownership.  The columns of X are cut into 32-column tiles; tile t
// belongs to team t % 8 for every factor, and a workgroup's team is the XCD it
// runs on, read from the hardware (XCC_ID).  Each output tile of factor f
// needs only the same tile of factor f-1's output, so a team reads
// activations that its own XCD wrote: the intermediates stay in that XCD's L2
// and only the blocks (and the chain's input and output) cross the fabric.
//
// Work.  Per factor a team has nbr x nt (block row, tile) pairs, one per wave;
// a workgroup of WAVES waves claims WAVES consecutive pairs at a time from the
// team's claim counter, and counts finished claims on the team's done counter.
// A workgroup processes a pair of factor f once the team's done counter of
// factor f-1 covers all of factor f-1's items; it claims its first item of
// factor f+1, and loads that item's blocks, before it waits, so only the
// activation loads sit behind the wait.  Work is only claimed by running
// workgroups, so no wait depends on another workgroup becoming resident; a
// workgroup that arrives late finds the work claimed and does not wait.
//
// Publication.  Activations written inside the launch are stored and loaded
// at device scope (sc1 buffer stores / loads: they miss the CU's L1, whose
// lines another CU's writes do not invalidate), and every wave drains its
// stores (s_waitcnt vmcnt(0)) before its workgroup's done increment; the
// counters are agent-scope atomics and the polls agent-scope atomic loads.
// No cache-wide write-back or invalidate is needed: measured on MI355X, an
// L1 invalidate of workgroup scope (buffer_inv sc0) leaves stale lines, and
// an agent-scope acquire (buffer_inv sc1 in every workgroup) costs ~30 us per
// factor, while device-scope loads of just the activations cost nothing
// measurable.  Blocks, the chain's input X and its output Y use plain
// accesses (they are not written inside the launch before being read).
//
// Termination.  Every wait is bounded (spin_max polls) and also ends when the
// control block's abort word is set; a wait that runs out sets it.  The last
// workgroup to finish (a global finish counter) checks that every team that
// owns a tile took part (code 2 otherwise), ORs the call's codes into the
// sticky report word and clears everything else, so every call starts from
// zeroed counters.  A non-zero report (1: a wait ran out, 2: a team missing)
// means an output was not valid; the host reads and clears it (on a
// process's first chain, or after every chain with FH_BSR_CHAIN_CHECK=1).
//
// Arithmetic: a pair's accumulation order (slot groups of 4, blocks, K halves,
// hi then lo plane) and epilogue are those of k_bsr_shared, so the chain and
// the factor-by-factor product give the same bits.
#include <algorithm>

#include "bsr_frag.hpp"
#include "fh_api_common.hpp"

namespace fh {

constexpr int BSR_CHAIN_MAX = 16;
// control block: abort word, finish counter, team mask, report word, then one
// claim and one done counter per (team, factor), each on its own 128-byte line
constexpr int CHAIN_LINE = 32;  // ints
constexpr int CHAIN_CTL_INTS = 4 * CHAIN_LINE + 2 * 8 * BSR_CHAIN_MAX * CHAIN_LINE;

struct BsrChainArgs {
  const int2* ell[BSR_CHAIN_MAX];   // (block index, block column) slots, W per block row
  const u16* ebdat[BSR_CHAIN_MAX];  // blocks in ELL order, operand layout
  int W[BSR_CHAIN_MAX];
  int J;
};

__device__ __forceinline__ int chain_rmw(int* p, int v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The first slot group of one (block row, tile) pair: its block columns and
// blocks, loaded before the wait for the previous factor (blocks do not
// depend on it).
template <bool SPLIT>
struct PairHead {
  bf16x8 a[4][SPLIT ? 2 : 1][2];
  int bcol[4];
};

template <bool SPLIT>
__device__ __forceinline__ void pair_head(const int2* __restrict__ ell, const u16* __restrict__ ebdat, int W, int br,
                                          int g, int lane, PairHead<SPLIT>& h) {
  constexpr int BLK = 1024;
  constexpr int PLANES = SPLIT ? 2 : 1;
  const int4* slots = reinterpret_cast<const int4*>(ell + (size_t)br * W);
  const int4 s01 = slots[g / 2], s23 = slots[g / 2 + 1];
  h.bcol[0] = s01.y;
  h.bcol[1] = s01.w;
  h.bcol[2] = s23.y;
  h.bcol[3] = s23.w;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const u16* blk = ebdat + ((size_t)br * W + g + q) * BLK * PLANES + lane * 8;
#pragma unroll
    for (int pl = 0; pl < PLANES; ++pl)
#pragma unroll
      for (int k = 0; k < 2; ++k) h.a[q][pl][k] = *reinterpret_cast<const bf16x8*>(blk + pl * BLK + k * 512);
  }
}

// the rest of the pair: activation loads, MFMAs (slot groups in order; the
// first from `head`), epilogue
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
constexpr int AUX_SC1 = 16;          // cache-policy operand of the buffer intrinsics: sc1, device scope
constexpr int RSRC_WORD3 = 0x00020000;  // buffer resource word 3 for raw dword access (gfx9)

template <bool SPLIT>
__device__ __forceinline__ void pair_finish(const int2* __restrict__ ell, const u16* __restrict__ ebdat, int W, int br,
                                            int tile, int G, int p, int n, const u16* __restrict__ X,
                                            u16* __restrict__ Y, int lane, PairHead<SPLIT>& head, bool IN_TILED,
                                            bool OUT_TILED) {
  constexpr int PLANES = SPLIT ? 2 : 1;
  const int kh = lane >> 5;
  const int jc = tile * 32 + (lane & 31);
  const bool jok = jc < p;
  const int j = jok ? jc : 0;
  // tiled activations are the intermediates written in this launch: device scope
  const int tiled_bytes = n * G * 32 * 2;
  const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc((void*)X, 0, tiled_bytes, RSRC_WORD3);
  f32x16 acc = {};
  for (int g = 0; g < W; g += 4) {
    if (g > 0) pair_head<SPLIT>(ell, ebdat, W, br, g, lane, head);
    bf16x8 x[4][2];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (IN_TILED) {
          const int off = (int)((frag_tile(head.bcol[q], G, tile) + (k * 64 + lane) * 8) * 2);
          x[q][k] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rx, off, 0, AUX_SC1));
        } else {
          x[q][k] = *reinterpret_cast<const bf16x8*>(X + (size_t)j * n + head.bcol[q] * 32 + k * 16 + kh * 8);
        }
      }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(head.a[q][0][k], x[q][k], acc, 0, 0, 0);
        if (SPLIT) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(head.a[q][PLANES - 1][k], x[q][k], acc, 0, 0, 0);
      }
  }
  if (!jok) return;
  const __amdgpu_buffer_rsrc_t ry = __builtin_amdgcn_make_buffer_rsrc((void*)Y, 0, tiled_bytes, RSRC_WORD3);
#pragma unroll
  for (int gg = 0; gg < 4; ++gg) {
    if (OUT_TILED) {
      // bsr_store4's epilogue with alpha = 1, beta = 0, stored at device scope
      const u32x2 w = {f2bf2(acc[4 * gg], acc[4 * gg + 1]), f2bf2(acc[4 * gg + 2], acc[4 * gg + 3])};
      const int off = (int)((frag_tile(br, G, tile) + frag_store_off(gg, kh, lane & 31)) * 2);
      __builtin_amdgcn_raw_buffer_store_b64(w, ry, off, 0, AUX_SC1);
    } else {
      bsr_store4(Y + (size_t)j * n + br * 32 + 8 * gg + 4 * kh, acc[4 * gg], acc[4 * gg + 1], acc[4 * gg + 2],
                 acc[4 * gg + 3], 1.0f, 0.0f);
    }
  }
}

template <bool SPLIT, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void k_bsr_chain(BsrChainArgs A, int nbr, int p, const u16* __restrict__ X,
                                                          u16* __restrict__ Y, u16* __restrict__ T0,
                                                          u16* __restrict__ T1, int* __restrict__ ctl, int spin_max,
                                                          int zero, int wait_extra) {
  __shared__ int s_item, s_next, s_stop;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  unsigned xreg;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(xreg));
  const int team = (int)(xreg & 7u);
  const int J = A.J, G = (p + 31) >> 5, n = nbr * 32;
  const int nt = G > team ? (G - team + 7) / 8 : 0;  // tiles team, team + 8, ...
  const int npairs = nbr * nt;
  const int items = (npairs + WAVES - 1) / WAVES;
  int* const err = ctl;
  int* const fin = ctl + CHAIN_LINE;
  int* const mask = ctl + 2 * CHAIN_LINE;
  int* const report = ctl + 3 * CHAIN_LINE;
  int* const claim = ctl + 4 * CHAIN_LINE + (team * J) * CHAIN_LINE;
  int* const done = ctl + 4 * CHAIN_LINE + (8 * J + team * J) * CHAIN_LINE;
  if (threadIdx.x == 0) {
    s_stop = 0;
    s_next = -1;
    __hip_atomic_fetch_or(mask, 1 << team, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_item = chain_rmw(claim, 1);
  }
  __syncthreads();
  int item = s_item;
  for (int f = 0; f < J; ++f) {
    const u16* in = f == 0 ? X : (((f - 1) & 1) ? T1 : T0);
    u16* out = f == J - 1 ? Y : ((f & 1) ? T1 : T0);
    bool waited = f == 0;  // factor f's activations are ready once done[f-1] covers factor f-1
    while (item < items) {
      const int q = item * WAVES + wave;
      const bool has = q < npairs;
      const int br = has ? q / nt : 0, tile = has ? team + 8 * (q % nt) : 0;
      PairHead<SPLIT> head;
      if (has) pair_head<SPLIT>(A.ell[f], A.ebdat[f], A.W[f], br, 0, lane, head);
      if (!waited) {
        if (threadIdx.x == 0) {
          int k = 0;
          bool ok = false;
          for (;;) {
            if (__hip_atomic_load(done + (f - 1) * CHAIN_LINE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
                items + wait_extra) {
              ok = true;
              break;
            }
            if ((k & 15) == 15 && chain_rmw(err, zero) != 0) break;
            if (++k >= spin_max) break;
            __builtin_amdgcn_s_sleep(1);
          }
          if (!ok) {
            __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_stop = 1;
          }
        }
        __syncthreads();
        if (s_stop) break;
        waited = true;
      }
      // intermediates are tiled; the chain's input and output column-major
      if (has)
        pair_finish<SPLIT>(A.ell[f], A.ebdat[f], A.W[f], br, tile, G, p, n, in, out, lane, head, f > 0, f < J - 1);
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wave's stores are complete
      __syncthreads();  // every wave's stores are complete; every wave has read s_item
      if (threadIdx.x == 0) {
        // three independent atomics in flight together: this item is done,
        // the next item of factor f, and (once) the first item of factor f+1
        __hip_atomic_fetch_add(done + f * CHAIN_LINE, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int a = chain_rmw(claim + f * CHAIN_LINE, 1);
        int b = s_next;
        if (f + 1 < J && b < 0) b = chain_rmw(claim + (f + 1) * CHAIN_LINE, 1);
        s_item = a;
        s_next = b;
      }
      __syncthreads();
      item = s_item;
    }
    if (s_stop || f == J - 1) break;
    __syncthreads();  // every wave has read s_item
    if (threadIdx.x == 0) {
      int b = s_next;
      if (b < 0) b = chain_rmw(claim + (f + 1) * CHAIN_LINE, 1);  // no item in factor f: claim now
      s_item = b;
      s_next = -1;
    }
    __syncthreads();
    item = s_item;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // the last workgroup: check that every tile-owning team ran, report, clear
    if (__hip_atomic_fetch_add(fin, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == (int)gridDim.x - 1) {
      const int need = G >= 8 ? 0xff : (1 << G) - 1;
      const int seen = chain_rmw(mask, zero);
      const int code = chain_rmw(err, zero) | ((seen & need) != need ? 2 : 0);
      if (code) __hip_atomic_fetch_or(report, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(err, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (int t = 0; t < 8 * J; ++t) {
        __hip_atomic_store(ctl + 4 * CHAIN_LINE + t * CHAIN_LINE, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(ctl + 4 * CHAIN_LINE + (8 * J + t) * CHAIN_LINE, 0, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
      __hip_atomic_store(mask, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(fin, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

__global__ void k_chain_ctl_clear(int* ctl, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) ctl[i] = 0;
}

size_t bsr_chain_ctl_bytes() { return sizeof(int) * CHAIN_CTL_INTS; }

void bsr_chain_ctl_clear(int* ctl, hipStream_t st) {
  hipLaunchKernelGGL(k_chain_ctl_clear, dim3(1), dim3(256), 0, st, ctl, CHAIN_CTL_INTS);
  FH_LAUNCH_CHECK();
}

// Read (synchronously) and clear the report word of a control block.
int bsr_chain_report(int* ctl, hipStream_t st) {
  int code = 0;
  FH_HIP(hipMemcpyAsync(&code, ctl + 3 * CHAIN_LINE, sizeof(int), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  if (code) bsr_chain_ctl_clear(ctl, st);
  return code;
}

// Whether the one-launch chain is asked for: FH_BSR_CHAIN=1 or the FH_MUL_FUSED
// mul mode (Faust.optimize_time picks it when it times faster); FH_BSR_CHAIN=0,
// a per-factor mode or a BSR kernel mode turn it off.  Returns 0 (use it) or
// -1.  On MI355X at config 3a it matches six launches, so it is not the default.
int bsr_chain_sync_mode() {
  const int m = mul_mode();
  if (m == FH_MUL_PER_FACTOR || (m >= FH_MUL_BSR_KERNEL(0) && m <= FH_MUL_BSR_KERNEL(5))) return -1;
  const char* e = getenv("FH_BSR_CHAIN");
  if (e && e[0] == '0') return -1;
  if (e && e[0] == '1') return 0;
  return m == FH_MUL_FUSED ? 0 : -1;
}

bool bsr_chain_fits(const std::vector<DMatP>& f, char op) {
  const int J = (int)f.size();
  if (J < 2 || J > BSR_CHAIN_MAX) return false;
  const bool tr = op != 'N';
  const int n = f[0]->m;
  if (n % 32 != 0 || n <= 0) return false;
  for (auto& x : f) {
    if (x->kind != K_BSR || x->m != n || x->n != n || x->split != f[0]->split) return false;
    const int W = tr ? x->tell_w : x->ell_w;
    if (W <= 0 || W % 4 != 0) return false;
  }
  return true;
}

// Launch the chain: Y = op(F) X, X and Y column-major (ld = n), T0 / T1 tiled
// intermediates (n x 32 ceil(p/32) BF16 each), ctl the cleared control block.
void bsr_chain_launch(const std::vector<DMatP>& f, char op, const void* X, int p, void* Y, void* T0, void* T1,
                      int* ctl, hipStream_t st) {
  const int J = (int)f.size();
  const bool tr = op != 'N';
  BsrChainArgs A{};
  A.J = J;
  for (int i = 0; i < J; ++i) {
    const DMat& S = (op == 'N') ? *f[J - 1 - i] : *f[i];
    A.ell[i] = (tr ? S.tell : S.ell).as<int2>();
    A.ebdat[i] = (tr ? S.tebdat : S.ebdat).as<u16>();
    A.W[i] = tr ? S.tell_w : S.ell_w;
  }
  const int n = f[0]->m, nbr = n / 32;
  const int G = (p + 31) / 32;
  FH_CHECK(((uintptr_t)X & 15) == 0 && ((uintptr_t)Y & 7) == 0 && ((uintptr_t)T0 & 15) == 0 &&
               ((uintptr_t)T1 & 15) == 0,
           "BSR chain: misaligned activations");
  // 16 waves per CU (one workgroup of 16, or two of 8 for FP32 hi/lo blocks,
  // whose wave needs more registers): 32 (64) workgroups per team, each
  // team's pairs split over its workgroups (any more would only wait)
  const bool split = f[0]->split;
  const int waves = split ? 8 : 16;
  const int max_nt = (G + 7) / 8;
  const int items = (nbr * max_nt + waves - 1) / waves;
  const int per_team = std::max(1, std::min(512 / waves, items));
  const dim3 grid((unsigned)(8 * per_team));
  const char* e_spin = getenv("FH_BSR_CHAIN_SPIN");
  const int spin_max = e_spin ? std::max(1, atoi(e_spin)) : (1 << 16);
  // test hook: FH_BSR_CHAIN_TEST_UNREACHABLE=1 asks every wait for one more
  // item than exists, so the waits run out (the failure path, forced from the host)
  const char* e_unr = getenv("FH_BSR_CHAIN_TEST_UNREACHABLE");
  const int wait_extra = (e_unr && e_unr[0] == '1') ? 1 : 0;
#define FH_CHAIN(S, WV) \
  hipLaunchKernelGGL((k_bsr_chain<S, WV>), grid, dim3(WV * 64), 0, st, A, nbr, p, (const u16*)X, (u16*)Y, \
                     (u16*)T0, (u16*)T1, ctl, spin_max, 0, wait_extra)
  if (split)
    FH_CHAIN(true, 8);
  else
    FH_CHAIN(false, 16);
#undef FH_CHAIN
  FH_LAUNCH_CHECK();
}

}  // namespace fh
