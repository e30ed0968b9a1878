// k_mixed.hip -- FP32 / BF16 Faust x dense: BSR on MFMA and FP32 CSR SpMM.
//
// BSR (32x32 blocks, bm = bn = 32): Y = op(B) X with X, Y BF16, FP32
// accumulation in v_mfma_f32_32x32x16_bf16 (K = 16: two MFMAs per block and
// 32x32 output tile).  Operand lanes:
//   A: lane l holds block row (l % 32), 8 consecutive k at 8*(l / 32);
//   B: lane l holds X column (l % 32), the same 8 consecutive k;
// blocks are stored in that operand order (entry (i, k) of a block at
// ((k / 16)*64 + ((k / 8) % 2)*32 + i)*8 + k % 8), so the A operand of one K
// half is one contiguous 1 KiB load, and X column-major (or tiled, below)
// gives the B operand as one 16-byte load per lane.  FP32 block values are
// stored as a BF16 (hi, lo) pair and cost one extra MFMA per step (error
// ~2^-16 relative per value, below the BF16 rounding of the activations).
// Default kernel (k_bsr_shared): a workgroup owns one block-row x 128
// columns, one 32-column tile per wave; the row's blocks go through LDS once
// per workgroup and the four waves gather their X tiles.  The blocks are read
// from a copy stored in ELL order (slot q of block row br at block
// br * W + q), so the block loads do not wait for the slot loads that give
// the X block columns.  Consecutive
// workgroup ids walk the column groups first, so with 512 columns each XCD
// keeps one 128-column group (and half of the block rows) and its L2 serves
// the repeated gathers of that group.  Other forms (one or more tiles per
// wave without LDS, K split over two waves, a row-pointer form for rows too
// wide for the ELL slots) are selectable.
//
// FP32 CSR: intermediates are row-major inside the chain so that the
// gather of one X row segment is a contiguous 128-byte read.  Factors with
// rows of at most 16 entries are read through their ELL view (index/value
// pairs at r * w, no row pointer in front of the gathers).  A workgroup
// computes 32 rows x 32 columns (8 lanes x float4 per row); panels of 32
// columns cycle over the XCDs the same way.  The chain input is transposed
// once into row-major and the last factor writes column-major through LDS.
#include <algorithm>
#include <vector>

#include "bsr_frag.hpp"
#include "fh_api_common.hpp"

namespace fh {

static inline u16 host_f2bf(float f) {
  unsigned u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (u16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
static inline float host_bf2f(u16 h) {
  unsigned u = ((unsigned)h) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// Barrier that orders LDS only: the X loads already issued stay in flight
// across it, so that a wave's first MFMAs start when their own operands
// arrive instead of after every load of the round (__syncthreads waits for
// all of them).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <bool SPLIT, bool IN_TILED, bool OUT_TILED>
__global__ __launch_bounds__(256) void k_bsr_mfma(int nbr, int p, const int* __restrict__ brp,
                                                  const int* __restrict__ bci, const u16* __restrict__ bdat,
                                                  const u16* __restrict__ X, int ldx, u16* __restrict__ Y, int ldy,
                                                  float alpha, float beta, int npanels) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int panel = blockIdx.x % npanels, quad = blockIdx.x / npanels;
  const int br = quad * 4 + wave;
  if (br >= nbr) return;
  const int kh = lane >> 5;  // which 8-wide k group of the 16
  const int G = (p + 31) >> 5;
  int jc[2], gc[2];
  bool jok[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    jc[c] = panel * 64 + c * 32 + (lane & 31);
    jok[c] = jc[c] < p;
    gc[c] = (panel * 2 + c < G) ? panel * 2 + c : 0;
  }
  constexpr int BLK = 1024;  // entries per 32x32 block
  constexpr int PLANES = SPLIT ? 2 : 1;
  f32x16 acc[2] = {{}, {}};
  const int b0 = brp[br], b1 = brp[br + 1];
  for (int b = b0; b < b1; b += 4) {
    bf16x8 a[4][2][PLANES], x[4][2][2];
    const int nb = min(4, b1 - b);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < nb) {
        const int bc = bci[b + q];
        const u16* blk = bdat + (size_t)(b + q) * BLK * PLANES + lane * 8;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
          for (int s = 0; s < PLANES; ++s) a[q][h][s] = *reinterpret_cast<const bf16x8*>(blk + s * BLK + h * 512);
#pragma unroll
          for (int c = 0; c < 2; ++c) {
            const int j = jok[c] ? jc[c] : 0;
            const u16* src = IN_TILED ? X + frag_tile(bc, G, gc[c]) + (h * 64 + lane) * 8
                                      : X + (size_t)j * ldx + bc * 32 + h * 16 + kh * 8;
            x[q][h][c] = jok[c] ? *reinterpret_cast<const bf16x8*>(src) : bf16x8{};
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < nb) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int c = 0; c < 2; ++c) {
            acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[q][h][0], x[q][h][c], acc[c], 0, 0, 0);
            if (SPLIT) acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[q][h][PLANES - 1], x[q][h][c], acc[c], 0, 0, 0);
          }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if (!jok[c]) continue;
    const int j = jc[c];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      u16* yq = OUT_TILED ? Y + frag_tile(br, G, gc[c]) + frag_store_off(g, kh, lane & 31)
                          : Y + (size_t)j * ldy + br * 32 + 8 * g + 4 * kh;
      bsr_store4(yq, acc[c][4 * g], acc[c][4 * g + 1], acc[c][4 * g + 2], acc[c][4 * g + 3], alpha, beta);
    }
  }
}

// ELL form of the same product: a wave reads its block-row's slots
// (block index, block column) as one wave-uniform scalar load, so the X and
// block loads follow after a single index latency, and it covers NC 32-column
// tiles (NC = 4: 128 columns, four accumulators) so that more bytes are in
// flight behind that latency.  Workgroup = 4 block-rows x 32*NC columns.
constexpr int BSR_ELL_MAX = 16;

// Workgroup -> (column panel, quad of block rows).  split == 1: consecutive
// ids walk the panels (each XCD keeps fixed panels and re-reads every block);
// split == 2: the XCDs form 8/2 column groups x 2 halves of the block rows,
// so every block is read by half as many XCDs while each X panel slice still
// stays in one XCD's L2.
__device__ __forceinline__ void bsr_map(int bid, int P, int Q, int split, int& panel, int& quad) {
  if (split == 1) {
    panel = bid % P;
    quad = bid / P;
    return;
  }
  const int G = 8 / split, ppg = P / G;
  const int x = bid % 8, slot = bid / 8;
  panel = (x % G) * ppg + slot % ppg;
  quad = (x / G) * (Q / split) + slot / ppg;
}

template <class V>
__device__ __forceinline__ V ld_stream(const V* p, bool nt) {
  return nt ? __builtin_nontemporal_load(p) : *p;
}

template <bool SPLIT, bool IN_TILED, bool OUT_TILED, int NC, bool NT>
__global__ __launch_bounds__(256) void k_bsr_ell(int nbr, int p, int W, const int2* __restrict__ ell,
                                                 const u16* __restrict__ bdat, const u16* __restrict__ X, int ldx,
                                                 u16* __restrict__ Y, int ldy, float alpha, float beta,
                                                 int npanels, int nquads, int xsplit) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int panel, quad;
  bsr_map(blockIdx.x, npanels, nquads, xsplit, panel, quad);
  const int br = quad * 4 + wave;
  if (br >= nbr) return;
  const int kh = lane >> 5;
  const int G = (p + 31) >> 5;
  int jc[NC], gc[NC];
  bool jok[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    jc[c] = panel * 32 * NC + c * 32 + (lane & 31);
    jok[c] = jc[c] < p;
    gc[c] = (panel * NC + c < G) ? panel * NC + c : 0;  // groups past p read group 0
  }
  constexpr int BLK = 1024;
  constexpr int PLANES = SPLIT ? 2 : 1;
  f32x16 acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = f32x16{};
  // W is a multiple of 4 and empty slots point at an all-zero block, so every
  // load below is unconditional: the whole group's loads are in flight
  // together.  Columns beyond p read column 0 (any finite data): they only
  // feed accumulator columns that are never stored.
  const int4* slots = reinterpret_cast<const int4*>(ell + (size_t)br * W);
  int j[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) j[c] = jok[c] ? jc[c] : 0;
  for (int g = 0; g < W; g += 4) {
    const int4 s01 = slots[g / 2], s23 = slots[g / 2 + 1];
    const int bi[4] = {s01.x, s01.z, s23.x, s23.z};
    const int bcol[4] = {s01.y, s01.w, s23.y, s23.w};
    bf16x8 a[4][2][PLANES], x[4][2][NC];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const u16* blk = bdat + (size_t)bi[q] * BLK * PLANES + lane * 8;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int pl = 0; pl < PLANES; ++pl)
          a[q][h][pl] = ld_stream(reinterpret_cast<const bf16x8*>(blk + pl * BLK + h * 512), NT);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const u16* src = IN_TILED ? X + frag_tile(bcol[q], G, gc[c]) + (h * 64 + lane) * 8
                                    : X + (size_t)j[c] * ldx + bcol[q] * 32 + h * 16 + kh * 8;
          x[q][h][c] = *reinterpret_cast<const bf16x8*>(src);
        }
      }
    }
    // keep the scheduler from interleaving the loads with the MFMAs (it
    // would, to save registers, and leave only a few loads in flight)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[q][h][0], x[q][h][c], acc[c], 0, 0, 0);
          if (SPLIT)
            acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[q][h][PLANES - 1], x[q][h][c], acc[c], 0, 0, 0);
        }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (!jok[c]) continue;
    const int j = jc[c];
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
      u16* yq = OUT_TILED ? Y + frag_tile(br, G, gc[c]) + frag_store_off(gg, kh, lane & 31)
                          : Y + (size_t)j * ldy + br * 32 + 8 * gg + 4 * kh;
      bsr_store4<NT>(yq, acc[c][4 * gg], acc[c][4 * gg + 1], acc[c][4 * gg + 2], acc[c][4 * gg + 3], alpha, beta);
    }
  }
}

// K-split ELL form: two waves share one block-row x 32 columns, each taking
// one 16-wide K half of every block (half the loads and MFMAs per wave, twice
// the waves in flight).  The two partial accumulators meet in LDS; each wave
// then finishes and stores half of the output rows (K half 0 + K half 1).
// Workgroup = 2 block-rows x 2 K halves x 32 columns.  A trailing odd block
// row is computed twice (clamped) and stored once, so every wave reaches the
// barrier.
template <bool SPLIT, bool IN_TILED, bool OUT_TILED>
__global__ __launch_bounds__(256) void k_bsr_ks(int nbr, int p, int W, const int2* __restrict__ ell,
                                                const u16* __restrict__ bdat, const u16* __restrict__ X, int ldx,
                                                u16* __restrict__ Y, int ldy, float alpha, float beta,
                                                int npanels) {
  __shared__ f32x16 part[4][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int panel = blockIdx.x % npanels, pair = blockIdx.x / npanels;
  const int hh = wave & 1;
  const int br_raw = pair * 2 + (wave >> 1);
  const bool row_ok = br_raw < nbr;
  const int br = row_ok ? br_raw : nbr - 1;
  const int kh = lane >> 5;
  const int G = (p + 31) >> 5;  // == npanels
  const int jc = panel * 32 + (lane & 31);
  const bool jok = jc < p;
  const int j = jok ? jc : 0;
  constexpr int BLK = 1024;
  constexpr int PLANES = SPLIT ? 2 : 1;
  f32x16 acc = {};
  const int4* slots = reinterpret_cast<const int4*>(ell + (size_t)br * W);
  for (int g = 0; g < W; g += 4) {
    const int4 s01 = slots[g / 2], s23 = slots[g / 2 + 1];
    const int bi[4] = {s01.x, s01.z, s23.x, s23.z};
    const int bcol[4] = {s01.y, s01.w, s23.y, s23.w};
    bf16x8 a[4][PLANES], x[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const u16* blk = bdat + (size_t)bi[q] * BLK * PLANES + (hh * 64 + lane) * 8;
#pragma unroll
      for (int pl = 0; pl < PLANES; ++pl) a[q][pl] = *reinterpret_cast<const bf16x8*>(blk + pl * BLK);
      const u16* src = IN_TILED ? X + frag_tile(bcol[q], G, panel) + (hh * 64 + lane) * 8
                                : X + (size_t)j * ldx + bcol[q] * 32 + hh * 16 + kh * 8;
      x[q] = *reinterpret_cast<const bf16x8*>(src);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[q][0], x[q], acc, 0, 0, 0);
      if (SPLIT) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[q][PLANES - 1], x[q], acc, 0, 0, 0);
    }
  }
  part[wave][lane] = acc;
  __syncthreads();
  const f32x16 k0 = part[wave & ~1][lane], k1 = part[wave | 1][lane];
  if (!row_ok || !jok) return;
#pragma unroll
  for (int gi = 0; gi < 2; ++gi) {
    const int gg = 2 * hh + gi;
    u16* yq = OUT_TILED ? Y + frag_tile(br, G, panel) + frag_store_off(gg, kh, lane & 31)
                        : Y + (size_t)j * ldy + br * 32 + 8 * gg + 4 * kh;
    bsr_store4(yq, k0[4 * gg] + k1[4 * gg], k0[4 * gg + 1] + k1[4 * gg + 1], k0[4 * gg + 2] + k1[4 * gg + 2],
               k0[4 * gg + 3] + k1[4 * gg + 3], alpha, beta);
  }
}

// Shared-block ELL form: a workgroup owns one block-row x 4 column tiles of
// 32 (one per wave).  The row's blocks are loaded once per workgroup into
// LDS (wave q loads block q of each group of 4 slots) and every wave reads
// its A operands from there, so each CU pulls a block from the L2 once
// instead of once per column tile.  Waves whose tile lies past p still load
// and pass the barriers; they store nothing.
template <bool SPLIT, bool IN_TILED, bool OUT_TILED>
__global__ __launch_bounds__(256) void k_bsr_shared(int nbr, int p, int W, const int2* __restrict__ ell,
                                                    const u16* __restrict__ bdat, const u16* __restrict__ X,
                                                    int ldx, u16* __restrict__ Y, int ldy, float alpha, float beta,
                                                    int npanels, int pgroups) {
  constexpr int BLK = 1024;
  constexpr int PLANES = SPLIT ? 2 : 1;
  __shared__ bf16x8 sA[4][PLANES][2][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int pg = blockIdx.x % pgroups, br = blockIdx.x / pgroups;
  const int panel_raw = pg * 4 + wave;
  const bool pok = panel_raw < npanels;
  const int panel = pok ? panel_raw : npanels - 1;
  const int kh = lane >> 5;
  const int G = (p + 31) >> 5;  // == npanels
  const int jc = panel * 32 + (lane & 31);
  const bool jok = pok && jc < p;
  const int j = jc < p ? jc : 0;
  f32x16 acc = {};
  const int4* slots = reinterpret_cast<const int4*>(ell + (size_t)br * W);
  for (int g = 0; g < W; g += 4) {
    const int4 s01 = slots[g / 2], s23 = slots[g / 2 + 1];
    const int bcol[4] = {s01.y, s01.w, s23.y, s23.w};
    bf16x8 a_own[PLANES][2], x[4][2];
    // bdat is the ELL-ordered block copy: this wave's block of slot group g
    // needs no slot load
    const u16* blk = bdat + ((size_t)br * W + g + wave) * BLK * PLANES + lane * 8;
#pragma unroll
    for (int pl = 0; pl < PLANES; ++pl)
#pragma unroll
      for (int h = 0; h < 2; ++h) a_own[pl][h] = *reinterpret_cast<const bf16x8*>(blk + pl * BLK + h * 512);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const u16* src = IN_TILED ? X + frag_tile(bcol[q], G, panel) + (h * 64 + lane) * 8
                                  : X + (size_t)j * ldx + bcol[q] * 32 + h * 16 + kh * 8;
        x[q][h] = *reinterpret_cast<const bf16x8*>(src);
      }
    __builtin_amdgcn_sched_barrier(0);
    if (g > 0) lds_barrier();  // every wave is done reading the previous group's blocks
#pragma unroll
    for (int pl = 0; pl < PLANES; ++pl)
#pragma unroll
      for (int h = 0; h < 2; ++h) sA[wave][pl][h][lane] = a_own[pl][h];
    lds_barrier();
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sA[q][0][h][lane], x[q][h], acc, 0, 0, 0);
        if (SPLIT) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sA[q][PLANES - 1][h][lane], x[q][h], acc, 0, 0, 0);
      }
  }
  if (!jok) return;
#pragma unroll
  for (int gg = 0; gg < 4; ++gg) {
    u16* yq = OUT_TILED ? Y + frag_tile(br, G, panel) + frag_store_off(gg, kh, lane & 31)
                        : Y + (size_t)j * ldy + br * 32 + 8 * gg + 4 * kh;
    bsr_store4(yq, acc[4 * gg], acc[4 * gg + 1], acc[4 * gg + 2], acc[4 * gg + 3], alpha, beta);
  }
}

static void bsr_shared_launch(bool split, bool in_tiled, bool out_tiled, dim3 grid, hipStream_t st, int nbr, int p,
                              int W, const int2* ell, const u16* dat, const u16* X, int ldx, u16* Y, int ldy,
                              float alpha, float beta, int P, int PG) {
  const int sel = (split ? 4 : 0) | (in_tiled ? 2 : 0) | (out_tiled ? 1 : 0);
#define FH_SH(S, I, O)                                                                                           \
  hipLaunchKernelGGL((k_bsr_shared<S, I, O>), grid, dim3(256), 0, st, nbr, p, W, ell, dat, X, ldx, Y, ldy, alpha, \
                     beta, P, PG)
  switch (sel) {
    case 0: FH_SH(false, false, false); break;
    case 1: FH_SH(false, false, true); break;
    case 2: FH_SH(false, true, false); break;
    case 3: FH_SH(false, true, true); break;
    case 4: FH_SH(true, false, false); break;
    case 5: FH_SH(true, false, true); break;
    case 6: FH_SH(true, true, false); break;
    default: FH_SH(true, true, true); break;
  }
#undef FH_SH
}

static void bsr_ks_launch(bool split, bool in_tiled, bool out_tiled, dim3 grid, hipStream_t st, int nbr, int p,
                          int W, const int2* ell, const u16* dat, const u16* X, int ldx, u16* Y, int ldy,
                          float alpha, float beta, int P) {
  const int sel = (split ? 4 : 0) | (in_tiled ? 2 : 0) | (out_tiled ? 1 : 0);
#define FH_KS(S, I, O) \
  hipLaunchKernelGGL((k_bsr_ks<S, I, O>), grid, dim3(256), 0, st, nbr, p, W, ell, dat, X, ldx, Y, ldy, alpha, beta, P)
  switch (sel) {
    case 0: FH_KS(false, false, false); break;
    case 1: FH_KS(false, false, true); break;
    case 2: FH_KS(false, true, false); break;
    case 3: FH_KS(false, true, true); break;
    case 4: FH_KS(true, false, false); break;
    case 5: FH_KS(true, false, true); break;
    case 6: FH_KS(true, true, false); break;
    default: FH_KS(true, true, true); break;
  }
#undef FH_KS
}

// Which BSR kernel: FH_BSR_KERNEL=0 forces the row-pointer kernel, 1 / 2 / 4
// the ELL kernel with that many 32-column tiles per wave, 3 the K-split ELL
// kernel, 5 the shared-block kernel (the default).  Measured on 8192^2,
// 4 blocks per block row, 512 columns, per factor inside the chain: BF16
// blocks 5: 6.9 us (1: 7.5, 3: 8.4); FP32 hi/lo blocks 5: 7.9 us (1: 11.0,
// 3: 10.7).  A wave's time is mostly the latency of its slot load and then of
// its gathers; what the CUs pull from the L2 per output tile decides the rest.
static int bsr_kernel_choice(bool) {
  const int m = mul_mode();
  if (m >= FH_MUL_BSR_KERNEL(0) && m <= FH_MUL_BSR_KERNEL(5)) return m - FH_MUL_BSR_KERNEL(0);
  const char* e = getenv("FH_BSR_KERNEL");
  if (e && e[0] >= '0' && e[0] <= '5') return e[0] - '0';
  return 5;
}

template <bool S, bool I, bool O, int NC, bool NT>
static void bsr_ell_go(dim3 grid, hipStream_t st, int nbr, int p, int W, const int2* ell, const u16* dat,
                       const u16* X, int ldx, u16* Y, int ldy, float alpha, float beta, int P, int Q, int xs) {
  hipLaunchKernelGGL((k_bsr_ell<S, I, O, NC, NT>), grid, dim3(256), 0, st, nbr, p, W, ell, dat, X, ldx, Y, ldy,
                     alpha, beta, P, Q, xs);
}

template <bool S, bool I, bool O>
static void bsr_ell_nc(int nc, bool nt, dim3 grid, hipStream_t st, int nbr, int p, int W, const int2* ell,
                       const u16* dat, const u16* X, int ldx, u16* Y, int ldy, float alpha, float beta, int P, int Q,
                       int xs) {
  if (nc == 4) {
    if (nt) bsr_ell_go<S, I, O, 4, true>(grid, st, nbr, p, W, ell, dat, X, ldx, Y, ldy, alpha, beta, P, Q, xs);
    else bsr_ell_go<S, I, O, 4, false>(grid, st, nbr, p, W, ell, dat, X, ldx, Y, ldy, alpha, beta, P, Q, xs);
  } else if (nc == 1) {
    if (nt) bsr_ell_go<S, I, O, 1, true>(grid, st, nbr, p, W, ell, dat, X, ldx, Y, ldy, alpha, beta, P, Q, xs);
    else bsr_ell_go<S, I, O, 1, false>(grid, st, nbr, p, W, ell, dat, X, ldx, Y, ldy, alpha, beta, P, Q, xs);
  } else {
    if (nt) bsr_ell_go<S, I, O, 2, true>(grid, st, nbr, p, W, ell, dat, X, ldx, Y, ldy, alpha, beta, P, Q, xs);
    else bsr_ell_go<S, I, O, 2, false>(grid, st, nbr, p, W, ell, dat, X, ldx, Y, ldy, alpha, beta, P, Q, xs);
  }
}

static void bsr_ell_launch(bool split, bool in_tiled, bool out_tiled, int nc, bool nt, dim3 grid, hipStream_t st,
                           int nbr, int p, int W, const int2* ell, const u16* dat, const u16* X, int ldx, u16* Y,
                           int ldy, float alpha, float beta, int P, int Q, int xs) {
  const int sel = (split ? 4 : 0) | (in_tiled ? 2 : 0) | (out_tiled ? 1 : 0);
#define FH_ELL(S, I, O) bsr_ell_nc<S, I, O>(nc, nt, grid, st, nbr, p, W, ell, dat, X, ldx, Y, ldy, alpha, beta, P, Q, xs)
  switch (sel) {
    case 0: FH_ELL(false, false, false); break;
    case 1: FH_ELL(false, false, true); break;
    case 2: FH_ELL(false, true, false); break;
    case 3: FH_ELL(false, true, true); break;
    case 4: FH_ELL(true, false, false); break;
    case 5: FH_ELL(true, false, true); break;
    case 6: FH_ELL(true, true, false); break;
    default: FH_ELL(true, true, true); break;
  }
#undef FH_ELL
}

// in_tiled / out_tiled select the activation layouts (see above); ldx/ldy are
// only used for column-major operands.
static void bsr_mul(const DMat& B, char op, const void* X, int ldx, int p, void* Y, int ldy, float alpha,
                    float beta, hipStream_t st, bool in_tiled = false, bool out_tiled = false) {
  const bool tr = op != 'N';
  const int rows = tr ? B.n : B.m;
  const int nbr = rows / 32;
  if (!in_tiled) FH_CHECK(ldx % 8 == 0, "BSR MFMA path: ldx must be a multiple of 8");
  if (!out_tiled) FH_CHECK(ldy % 4 == 0, "BSR MFMA path: ldy must be a multiple of 4");
  FH_CHECK(((uintptr_t)X & 15) == 0 && ((uintptr_t)Y & 7) == 0, "BSR MFMA path: X must be 16-byte aligned");
  if (nbr == 0 || p == 0) return;
  const int quads = (nbr + 3) / 4;
  const u16* dat = tr ? B.tbdat.as<u16>() : B.bdat.as<u16>();
  const int W = tr ? B.tell_w : B.ell_w;
  const int choice = bsr_kernel_choice(B.split);
  if (W > 0 && choice == 5) {
    const int2* ell = (tr ? B.tell : B.ell).as<int2>();
    const int npanels = (p + 31) / 32, pgroups = (npanels + 3) / 4;
    const dim3 grid((unsigned)(pgroups * nbr));
    const u16* edat = tr ? B.tebdat.as<u16>() : B.ebdat.as<u16>();
    bsr_shared_launch(B.split, in_tiled, out_tiled, grid, st, nbr, p, W, ell, edat, (const u16*)X, ldx, (u16*)Y,
                      ldy, alpha, beta, npanels, pgroups);
    FH_LAUNCH_CHECK();
    return;
  }
  if (W > 0 && choice == 3) {
    const int2* ell = (tr ? B.tell : B.ell).as<int2>();
    const int npanels = (p + 31) / 32;
    const dim3 grid((unsigned)(npanels * ((nbr + 1) / 2)));
    bsr_ks_launch(B.split, in_tiled, out_tiled, grid, st, nbr, p, W, ell, dat, (const u16*)X, ldx, (u16*)Y, ldy,
                  alpha, beta, npanels);
    FH_LAUNCH_CHECK();
    return;
  }
  if (W > 0 && choice != 0) {
    const int2* ell = (tr ? B.tell : B.ell).as<int2>();
    const int nc = choice;
    const int npanels = (p + 32 * nc - 1) / (32 * nc);
    const dim3 grid((unsigned)(npanels * quads));
    const char* e_nt = getenv("FH_BSR_NT");
    const bool nt = e_nt && e_nt[0] == '1';
    int xsplit = 1;
    if (const char* e = getenv("FH_BSR_XSPLIT")) xsplit = (e[0] == '2') ? 2 : 1;
    if (xsplit == 2 && ((npanels * quads) % 8 != 0 || npanels % 4 != 0 || quads % 2 != 0)) xsplit = 1;
    bsr_ell_launch(B.split, in_tiled, out_tiled, nc, nt, grid, st, nbr, p, W, ell, dat, (const u16*)X, ldx,
                   (u16*)Y, ldy, alpha, beta, npanels, quads, xsplit);
    FH_LAUNCH_CHECK();
    return;
  }
  const int npanels = (p + 63) / 64;
  const int* rp = tr ? B.tbrp.as<int>() : B.brp.as<int>();
  const int* ci = tr ? B.tbci.as<int>() : B.bci.as<int>();
  dim3 grid((unsigned)(npanels * quads));
#define FH_BSR_LAUNCH(S, I, O)                                                                                \
  hipLaunchKernelGGL((k_bsr_mfma<S, I, O>), grid, dim3(256), 0, st, nbr, p, rp, ci, dat, (const u16*)X, ldx, \
                     (u16*)Y, ldy, alpha, beta, npanels)
  const int sel = (B.split ? 4 : 0) | (in_tiled ? 2 : 0) | (out_tiled ? 1 : 0);
  switch (sel) {
    case 0: FH_BSR_LAUNCH(false, false, false); break;
    case 1: FH_BSR_LAUNCH(false, false, true); break;
    case 2: FH_BSR_LAUNCH(false, true, false); break;
    case 3: FH_BSR_LAUNCH(false, true, true); break;
    case 4: FH_BSR_LAUNCH(true, false, false); break;
    case 5: FH_BSR_LAUNCH(true, false, true); break;
    case 6: FH_BSR_LAUNCH(true, true, false); break;
    default: FH_BSR_LAUNCH(true, true, true); break;
  }
#undef FH_BSR_LAUNCH
  FH_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- FP32 BSR, FP32 activations
// Y = alpha*op(B)*X + beta*Y with FP32 blocks and FP32 X / Y (column-major),
// the arithmetic of MatBSR<float>::multiply (faust_MatBSR.hpp:828-836) on
// v_mfma_f32_32x32x2_f32: a wave owns one block row x one 32-column tile and
// runs 16 K-steps per block; in step s lane (h, r) feeds A = B(r, 16h + s)
// and B = X(16h + s, tile column r), so the blocks (stored that way, 16
// consecutive floats per lane) and the X column segments are 4 float4 loads
// per lane.  Blocks in row order (b ascending, then step by step), FP32
// accumulation.
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_bsr_f32(int nbr, int p, const int* __restrict__ brp,
                                                 const int* __restrict__ bci, const float* __restrict__ blk,
                                                 const float* __restrict__ X, int ldx, float* __restrict__ Y, int ldy,
                                                 float alpha, float beta, int ntiles) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tile = blockIdx.x % ntiles, br = (blockIdx.x / ntiles) * 4 + wave;
  if (br >= nbr) return;
  const int h = lane >> 5, r = lane & 31;
  const int jc = tile * 32 + r;
  const bool jok = jc < p;
  const int j = jok ? jc : 0;  // columns past p read column 0; their outputs are not stored
  f32x16 acc = {};
  const int b0 = brp[br], b1 = brp[br + 1];
  for (int b = b0; b < b1; ++b) {
    const int bc = bci[b];
    const f32x4* ap = reinterpret_cast<const f32x4*>(blk + (size_t)b * 1024 + lane * 16);
    const f32x4* xp = reinterpret_cast<const f32x4*>(X + (size_t)j * ldx + bc * 32 + h * 16);
    f32x4 a[4], x[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = ap[q];
      x[q] = xp[q];
    }
#pragma unroll
    for (int s4 = 0; s4 < 16; ++s4)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[s4 >> 2][s4 & 3], x[s4 >> 2][s4 & 3], acc, 0, 0, 0);
  }
  if (!jok) return;
#pragma unroll
  for (int gg = 0; gg < 4; ++gg) {
    f32x4* yq = reinterpret_cast<f32x4*>(Y + (size_t)j * ldy + br * 32 + 8 * gg + 4 * h);
    f32x4 v = {alpha * acc[4 * gg], alpha * acc[4 * gg + 1], alpha * acc[4 * gg + 2], alpha * acc[4 * gg + 3]};
    if (beta != 0.0f) {
      const f32x4 old = *yq;
      v[0] = fmaf(beta, old[0], v[0]);
      v[1] = fmaf(beta, old[1], v[1]);
      v[2] = fmaf(beta, old[2], v[2]);
      v[3] = fmaf(beta, old[3], v[3]);
    }
    *yq = v;
  }
}

// Shared-block form: a workgroup owns one block row x 128 columns (one
// 32-column tile per wave, as k_bsr_shared does for BF16), the row's blocks
// (up to 4 at a time, 16 KiB) are staged once in LDS for the four waves, and
// each wave issues the X loads of all of them before its first MFMA.  The
// MFMA sequence of every output tile is k_bsr_f32's (blocks ascending, 16
// K-steps each), so the bits are the same.  Consecutive workgroups walk the
// column groups first (one group per XCD pair, as in k_bsr_shared).
constexpr int BF32_CH = 4;  // blocks staged per round
__global__ __launch_bounds__(256) void k_bsr_f32s(int nbr, int p, const int* __restrict__ brp,
                                                  const int* __restrict__ bci, const float* __restrict__ blk,
                                                  const float* __restrict__ X, int ldx, float* __restrict__ Y, int ldy,
                                                  float alpha, float beta, int ngroups) {
  __shared__ f32x4 sa[BF32_CH * 256];  // BF32_CH blocks, 1024 floats each
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, t = threadIdx.x;
  const int g = blockIdx.x % ngroups, br = blockIdx.x / ngroups;
  if (br >= nbr) return;
  const int h = lane >> 5, r = lane & 31;
  const int jc = g * 128 + wave * 32 + r;
  const bool jok = jc < p;
  const int j = jok ? jc : 0;  // columns past p read column 0; their outputs are not stored
  f32x16 acc = {};
  const int b0 = brp[br], b1 = brp[br + 1];
  for (int c0 = b0; c0 < b1; c0 += BF32_CH) {
    const int nb = min(BF32_CH, b1 - c0);
    if (c0 > b0) lds_barrier();  // the previous round's blocks are consumed
    if (nb == BF32_CH) {
      // a full round: the staging loads and this wave's X column segments of
      // all its blocks are issued together, before any LDS write or MFMA
      f32x4 stg[BF32_CH], x[BF32_CH][4];
#pragma unroll
      for (int q = 0; q < BF32_CH; ++q) stg[q] = reinterpret_cast<const f32x4*>(blk + (size_t)(c0 + q) * 1024)[t];
#pragma unroll
      for (int q = 0; q < BF32_CH; ++q) {
        const f32x4* xp = reinterpret_cast<const f32x4*>(X + (size_t)j * ldx + bci[c0 + q] * 32 + h * 16);
#pragma unroll
        for (int u = 0; u < 4; ++u) x[q][u] = xp[u];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < BF32_CH; ++q) sa[q * 256 + t] = stg[q];
      lds_barrier();
#pragma unroll
      for (int q = 0; q < BF32_CH; ++q) {
        f32x4 a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = sa[q * 256 + lane * 4 + u];
#pragma unroll
        for (int s4 = 0; s4 < 16; ++s4)
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[s4 >> 2][s4 & 3], x[q][s4 >> 2][s4 & 3], acc, 0, 0, 0);
      }
    } else {
      // the row's last, partial round: one block at a time
      for (int q = 0; q < nb; ++q) {
        if (q > 0) __syncthreads();
        sa[t] = reinterpret_cast<const f32x4*>(blk + (size_t)(c0 + q) * 1024)[t];
        const f32x4* xp = reinterpret_cast<const f32x4*>(X + (size_t)j * ldx + bci[c0 + q] * 32 + h * 16);
        f32x4 x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = xp[u];
        __syncthreads();
        f32x4 a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = sa[lane * 4 + u];
#pragma unroll
        for (int s4 = 0; s4 < 16; ++s4)
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[s4 >> 2][s4 & 3], x[s4 >> 2][s4 & 3], acc, 0, 0, 0);
      }
    }
  }
  if (!jok) return;
#pragma unroll
  for (int gg = 0; gg < 4; ++gg) {
    f32x4* yq = reinterpret_cast<f32x4*>(Y + (size_t)j * ldy + br * 32 + 8 * gg + 4 * h);
    f32x4 v = {alpha * acc[4 * gg], alpha * acc[4 * gg + 1], alpha * acc[4 * gg + 2], alpha * acc[4 * gg + 3]};
    if (beta != 0.0f) {
      const f32x4 old = *yq;
      v[0] = fmaf(beta, old[0], v[0]);
      v[1] = fmaf(beta, old[1], v[1]);
      v[2] = fmaf(beta, old[2], v[2]);
      v[3] = fmaf(beta, old[3], v[3]);
    }
    *yq = v;
  }
}

void bsr_mul_f32(const DMat& B, char op, const float* X, int ldx, int p, float* Y, int ldy, float alpha, float beta,
                 hipStream_t st) {
  FH_CHECK(B.kind == K_BSR && B.dtype == FH_F32 && B.f32dat.p, "FP32 activations need an FP32 BSR factor");
  const bool tr = op != 'N';
  const int rows = tr ? B.n : B.m;
  const int nbr = rows / 32;
  FH_CHECK(ldx % 4 == 0 && ldy % 4 == 0 && ((uintptr_t)X & 15) == 0 && ((uintptr_t)Y & 15) == 0,
           "FP32 BSR path: X and Y must be 16-byte aligned with leading dimensions that are multiples of 4");
  if (nbr == 0 || p == 0) return;
  const int* brp = (tr ? B.tbrp : B.brp).as<int>();
  const int* bci = (tr ? B.tbci : B.bci).as<int>();
  const float* blk = (tr ? B.tf32dat : B.f32dat).as<float>();
  // FH_BSR_F32_KERNEL=0: one wave per (block row, tile), blocks from global memory
  const char* e = getenv("FH_BSR_F32_KERNEL");
  if (e && e[0] == '0') {
    const int ntiles = (p + 31) / 32;
    const dim3 grid((unsigned)(ntiles * ((nbr + 3) / 4)));
    hipLaunchKernelGGL(k_bsr_f32, grid, dim3(256), 0, st, nbr, p, brp, bci, blk, X, ldx, Y, ldy, alpha, beta, ntiles);
  } else {
    const int ngroups = (p + 127) / 128;
    hipLaunchKernelGGL(k_bsr_f32s, dim3((unsigned)(ngroups * nbr)), dim3(256), 0, st, nbr, p, brp, bci, blk, X, ldx,
                       Y, ldy, alpha, beta, ngroups);
  }
  FH_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- FP32 CSR
// A row's batch of NB (<= 16) index/value pairs reaches its 8 lanes with two
// loads per lane: lane s of the row's group loads entries s and 8 + s (entry
// indices clamped to the row's last entry), and ds_swizzle hands entry q to
// every lane of the group from lane q % 8.  All 8 lanes of a group must be
// active (a group is one row).
template <int Q, int NB>
struct CsrBcast {
  static __device__ __forceinline__ void go(int ia, int ib, float fa, float fb, int* cc, float* ww) {
    constexpr int PAT = 0x18 | ((Q & 7) << 5);  // lane (l & ~7) | (Q % 8) within each 32
    cc[Q] = __builtin_amdgcn_ds_swizzle(Q < 8 ? ia : ib, PAT);
    ww[Q] = __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(Q < 8 ? fa : fb), PAT));
    CsrBcast<Q + 1, NB>::go(ia, ib, fa, fb, cc, ww);
  }
};
template <int NB>
struct CsrBcast<NB, NB> {
  static __device__ __forceinline__ void go(int, int, float, float, int*, float*) {}
};
template <int NB>
__device__ __forceinline__ void csr_batch_idx(const int* __restrict__ ci, const float* __restrict__ v, int base,
                                              int nb, int s, int (&cc)[NB], float (&ww)[NB]) {
  static_assert(NB <= 16, "two entries per lane");
  const int ea = base + min(s, nb - 1);
  const int ia = ci[ea];
  const float fa = v[ea];
  int ib = ia;
  float fb = fa;
  if (NB > 8) {
    const int eb = base + min(8 + s, nb - 1);
    ib = ci[eb];
    fb = v[eb];
  }
  CsrBcast<0, NB>::go(ia, ib, fa, fb, cc, ww);
}

// Row-major gather SpMM: Y[r, c] = alpha * sum_e v[e] * X[ci[e], c] + beta*Y.
// OUT_COL: write Y column-major (ld = ldy) through an LDS transpose.
// ELL: ci / v are the ELL view (NB slots per row, column -1 in unused
// slots; rp unused), so a row's pairs load without the row pointer first;
// the terms and their order are those of the CSR form.
template <bool OUT_COL, int NB, bool ELL = false>
__global__ __launch_bounds__(256) void k_csr_rm_f32(int rows, int p, const int* __restrict__ rp,
                                                    const int* __restrict__ ci, const float* __restrict__ v,
                                                    const float* __restrict__ X, int ldx, float* __restrict__ Y,
                                                    int ldy, float alpha, float beta, int npanels) {
  __shared__ float tile[32][33];
  const int t = threadIdx.x;
  const int panel = blockIdx.x % npanels, rblk = blockIdx.x / npanels;
  const int r = rblk * 32 + (t >> 3);
  const int c0 = panel * 32 + (t & 7) * 4;
  const int cg = c0 < p ? c0 : 0;  // lanes past p gather column 0 and store nothing
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  if (ELL && r < rows) {
    int cc[NB];
    float ww[NB];
    csr_batch_idx<NB>(ci, v, r * NB, NB, t & 7, cc, ww);
    float4 xx[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) xx[q] = *reinterpret_cast<const float4*>(X + (size_t)max(cc[q], 0) * ldx + cg);
    __builtin_amdgcn_sched_barrier(0);  // all gathers in flight before the first FMA
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const bool on = cc[q] >= 0;  // unused slots are not accumulated
      acc.x = on ? fmaf(ww[q], xx[q].x, acc.x) : acc.x;
      acc.y = on ? fmaf(ww[q], xx[q].y, acc.y) : acc.y;
      acc.z = on ? fmaf(ww[q], xx[q].z, acc.z) : acc.z;
      acc.w = on ? fmaf(ww[q], xx[q].w, acc.w) : acc.w;
    }
  }
  if (!ELL && r < rows) {  // the same for the row's 8 lanes, as the broadcast needs
    const int e0 = rp[r], e1 = rp[r + 1];
    // batches of up to NB nonzeros (NB: the widest row rounded up to 4, at
    // most 16): the index/value pairs, then all row gathers, then the FMAs,
    // so NB gathers are in flight per thread.  The loads are unconditional
    // (indices clamped to the row's last entry; the extra terms are not
    // accumulated): a conditional load would become a branch that waits for
    // the loads before it.
    for (int base = e0; base < e1; base += NB) {
      const int nb = min(NB, e1 - base);
      int cc[NB];
      float ww[NB];
      csr_batch_idx<NB>(ci, v, base, nb, t & 7, cc, ww);
      float4 xx[NB];
#pragma unroll
      for (int q = 0; q < NB; ++q) xx[q] = *reinterpret_cast<const float4*>(X + (size_t)cc[q] * ldx + cg);
      __builtin_amdgcn_sched_barrier(0);  // all gathers in flight before the first FMA
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const bool on = q < nb;  // terms past the row end are not accumulated
        acc.x = on ? fmaf(ww[q], xx[q].x, acc.x) : acc.x;
        acc.y = on ? fmaf(ww[q], xx[q].y, acc.y) : acc.y;
        acc.z = on ? fmaf(ww[q], xx[q].z, acc.z) : acc.z;
        acc.w = on ? fmaf(ww[q], xx[q].w, acc.w) : acc.w;
      }
    }
  }
  if (!OUT_COL) {
    if (r < rows && c0 < p) {
      float4* y = reinterpret_cast<float4*>(Y + (size_t)r * ldy + c0);
      float4 o = make_float4(alpha * acc.x, alpha * acc.y, alpha * acc.z, alpha * acc.w);
      if (beta != 0.0f) {
        const float4 old = *y;
        o.x = fmaf(beta, old.x, o.x);
        o.y = fmaf(beta, old.y, o.y);
        o.z = fmaf(beta, old.z, o.z);
        o.w = fmaf(beta, old.w, o.w);
      }
      *y = o;
    }
    return;
  }
  const int lr = t >> 3, lc = (t & 7) * 4;
  tile[lr][lc + 0] = acc.x;
  tile[lr][lc + 1] = acc.y;
  tile[lr][lc + 2] = acc.z;
  tile[lr][lc + 3] = acc.w;
  __syncthreads();
  // thread -> column (t >> 3), 4 consecutive rows (t & 7) * 4
  const int cc = t >> 3, rr = (t & 7) * 4;
  const int gc = panel * 32 + cc;
  if (gc >= p) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int gr = rblk * 32 + rr + q;
    if (gr < rows) {
      float o = alpha * tile[rr + q][cc];
      float* y = Y + (size_t)gc * ldy + gr;
      if (beta != 0.0f) o = fmaf(beta, *y, o);
      *y = o;
    }
  }
}

// Two rows per thread (workgroup = 64 rows x 32 columns): the index loads
// and the gathers of both rows are issued together, twice the bytes in flight
// behind each latency of the row-pointer -> index -> gather chain.
template <bool OUT_COL>
__global__ __launch_bounds__(256) void k_csr_rm2_f32(int rows, int p, const int* __restrict__ rp,
                                                     const int* __restrict__ ci, const float* __restrict__ v,
                                                     const float* __restrict__ X, int ldx, float* __restrict__ Y,
                                                     int ldy, float alpha, float beta, int npanels) {
  __shared__ float tile[64][33];
  constexpr int NB = 8;
  const int t = threadIdx.x;
  const int panel = blockIdx.x % npanels, rblk = blockIdx.x / npanels;
  const int c0 = panel * 32 + (t & 7) * 4;
  int r[2], e0[2], e1[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    r[i] = rblk * 64 + 32 * i + (t >> 3);
    const bool ok = r[i] < rows && c0 < p;
    e0[i] = ok ? rp[r[i]] : 0;
    e1[i] = ok ? rp[r[i] + 1] : 0;
  }
  float4 acc[2] = {make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f)};
  const int len = max(e1[0] - e0[0], e1[1] - e0[1]);
  for (int base = 0; base < len; base += NB) {
    int cc[2][NB];
    float ww[2][NB];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int e = e0[i] + base + q;
        const bool ok = e < e1[i];
        cc[i][q] = ok ? ci[e] : -1;
        ww[i][q] = ok ? v[e] : 0.0f;
      }
    float4 xx[2][NB];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int q = 0; q < NB; ++q)
        xx[i][q] = cc[i][q] >= 0 ? *reinterpret_cast<const float4*>(X + (size_t)cc[i][q] * ldx + c0)
                                 : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        acc[i].x = fmaf(ww[i][q], xx[i][q].x, acc[i].x);
        acc[i].y = fmaf(ww[i][q], xx[i][q].y, acc[i].y);
        acc[i].z = fmaf(ww[i][q], xx[i][q].z, acc[i].z);
        acc[i].w = fmaf(ww[i][q], xx[i][q].w, acc[i].w);
      }
  }
  if (!OUT_COL) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      if (r[i] < rows && c0 < p) {
        float4* y = reinterpret_cast<float4*>(Y + (size_t)r[i] * ldy + c0);
        float4 o = make_float4(alpha * acc[i].x, alpha * acc[i].y, alpha * acc[i].z, alpha * acc[i].w);
        if (beta != 0.0f) {
          const float4 old = *y;
          o.x = fmaf(beta, old.x, o.x);
          o.y = fmaf(beta, old.y, o.y);
          o.z = fmaf(beta, old.z, o.z);
          o.w = fmaf(beta, old.w, o.w);
        }
        *y = o;
      }
    }
    return;
  }
  const int lc = (t & 7) * 4;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int lr = 32 * i + (t >> 3);
    tile[lr][lc + 0] = acc[i].x;
    tile[lr][lc + 1] = acc[i].y;
    tile[lr][lc + 2] = acc[i].z;
    tile[lr][lc + 3] = acc[i].w;
  }
  __syncthreads();
  // thread -> column (t >> 3), 8 consecutive rows (t & 7) * 8
  const int cc = t >> 3, rr = (t & 7) * 8;
  const int gc = panel * 32 + cc;
  if (gc >= p) return;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int gr = rblk * 64 + rr + q;
    if (gr < rows) {
      float o = alpha * tile[rr + q][cc];
      float* y = Y + (size_t)gc * ldy + gr;
      if (beta != 0.0f) o = fmaf(beta, *y, o);
      *y = o;
    }
  }
}

// FH_CSR_RPT=2 selects the two-rows-per-thread kernel (measured slower on
// 16384^2, 12 nnz/row, 256 columns: 18.9 vs 17.5 us per factor)
static int csr_rows_per_thread() {
  if (mul_mode() == FH_MUL_CSR_ROWS2) return 2;
  if (mul_mode() == FH_MUL_CSR_ROWS1) return 1;
  const char* e = getenv("FH_CSR_RPT");
  return (e && e[0] == '2') ? 2 : 1;
}

// column-major (rows x p, ld) -> row-major (rows x ldr), zero-padding columns p..ldr
__global__ __launch_bounds__(256) void k_col2row_f32(int rows, int p, const float* __restrict__ X, int ldx,
                                                     float* __restrict__ R, int ldr) {
  __shared__ float tile[32][33];
  const int t = threadIdx.x;
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  // read: thread -> column c0 + (t >> 3), rows r0 + (t & 7)*4 .. +3
  const int cc = t >> 3, rr = (t & 7) * 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int gr = r0 + rr + q, gc = c0 + cc;
    tile[rr + q][cc] = (gr < rows && gc < p) ? X[(size_t)gc * ldx + gr] : 0.0f;
  }
  __syncthreads();
  const int lr = t >> 3, lc = (t & 7) * 4;
  const int gr = r0 + lr;
  if (gr >= rows) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int gc = c0 + lc + q;
    if (gc < ldr) R[(size_t)gr * ldr + gc] = tile[lr][lc + q];
  }
}

static void csr_f32_mul_rm(const DMat& S, char op, const float* X, int ldx, int p, float* Y, int ldy, bool out_col,
                           float alpha, float beta, hipStream_t st) {
  const bool tr = op != 'N';
  const int rows = tr ? S.n : S.m;
  if (rows == 0 || p == 0) return;
  const int* rp = tr ? S.cp.as<int>() : S.rp.as<int>();
  const int* ci = tr ? S.ri.as<int>() : S.ci.as<int>();
  const float* v = tr ? S.cv.as<float>() : S.v.as<float>();
  const int npanels = (p + 31) / 32;
  if (csr_rows_per_thread() == 2) {
    const dim3 grid2((unsigned)(npanels * ((rows + 63) / 64)));
    if (out_col)
      hipLaunchKernelGGL(k_csr_rm2_f32<true>, grid2, dim3(256), 0, st, rows, p, rp, ci, v, X, ldx, Y, ldy, alpha,
                         beta, npanels);
    else
      hipLaunchKernelGGL(k_csr_rm2_f32<false>, grid2, dim3(256), 0, st, rows, p, rp, ci, v, X, ldx, Y, ldy, alpha,
                         beta, npanels);
    FH_LAUNCH_CHECK();
    return;
  }
  dim3 grid((unsigned)(npanels * ((rows + 31) / 32)));
  const int ew = tr ? S.csr_tell_w : S.csr_ell_w;
  const char* e_ell = getenv("FH_CSR_ELL");
  if (ew > 0 && !(e_ell && e_ell[0] == '0')) {
    // ELL view: every row's pairs at r * ew, no row pointer in the chain
    const int* eci = tr ? S.csr_tell_ci.as<int>() : S.csr_ell_ci.as<int>();
    const float* ev = tr ? S.csr_tell_v.as<float>() : S.csr_ell_v.as<float>();
#define FH_CSR_ELL_GO(OC, NB)                                                                                       \
  hipLaunchKernelGGL((k_csr_rm_f32<OC, NB, true>), grid, dim3(256), 0, st, rows, p, nullptr, eci, ev, X, ldx, Y, ldy, \
                     alpha, beta, npanels)
    switch (ew * 2 + (out_col ? 1 : 0)) {
      case 8: FH_CSR_ELL_GO(false, 4); break;
      case 9: FH_CSR_ELL_GO(true, 4); break;
      case 16: FH_CSR_ELL_GO(false, 8); break;
      case 17: FH_CSR_ELL_GO(true, 8); break;
      case 24: FH_CSR_ELL_GO(false, 12); break;
      case 25: FH_CSR_ELL_GO(true, 12); break;
      case 32: FH_CSR_ELL_GO(false, 16); break;
      default: FH_CSR_ELL_GO(true, 16); break;
    }
#undef FH_CSR_ELL_GO
    FH_LAUNCH_CHECK();
    return;
  }
  // gather batch: the widest row rounded up to 4 (16 when unknown or wider)
  const int widest = tr ? S.max_col : S.max_row;
  const int nb = (widest < 0 || widest > 12) ? 16 : std::max(4, (widest + 3) / 4 * 4);
#define FH_CSR_GO(OC, NB)                                                                                  \
  hipLaunchKernelGGL((k_csr_rm_f32<OC, NB>), grid, dim3(256), 0, st, rows, p, rp, ci, v, X, ldx, Y, ldy, alpha, \
                     beta, npanels)
  switch (nb * 2 + (out_col ? 1 : 0)) {
    case 8: FH_CSR_GO(false, 4); break;
    case 9: FH_CSR_GO(true, 4); break;
    case 16: FH_CSR_GO(false, 8); break;
    case 17: FH_CSR_GO(true, 8); break;
    case 24: FH_CSR_GO(false, 12); break;
    case 25: FH_CSR_GO(true, 12); break;
    case 32: FH_CSR_GO(false, 16); break;
    default: FH_CSR_GO(true, 16); break;
  }
#undef FH_CSR_GO
  FH_LAUNCH_CHECK();
}

static void col2row(int rows, int p, const float* X, int ldx, float* R, int ldr, hipStream_t st) {
  if (rows == 0 || p == 0) return;
  dim3 grid((rows + 31) / 32, (ldr + 31) / 32);
  hipLaunchKernelGGL(k_col2row_f32, grid, dim3(256), 0, st, rows, p, X, ldx, R, ldr);
  FH_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- conversions
__global__ void k_bsr_to_dense(int nbr, const int* __restrict__ brp, const int* __restrict__ bci,
                               const u16* __restrict__ bdat, int split, int m, float* __restrict__ D) {
  const int br = blockIdx.x;
  for (int b = brp[br]; b < brp[br + 1]; ++b) {
    const u16* blk = bdat + (size_t)b * 1024 * (split ? 2 : 1);
    for (int e = threadIdx.x; e < 1024; e += blockDim.x) {
      // e in operand order: (k / 16, (k / 8) % 2, i, k % 8)
      const int k = ((e >> 9) << 4) | (((e >> 8) & 1) << 3) | (e & 7), i = (e >> 3) & 31;
      float v = bf2f(blk[e]);
      if (split) v += bf2f(blk[1024 + e]);
      D[(size_t)(bci[b] * 32 + k) * m + br * 32 + i] += v;
    }
    __syncthreads();
  }
}

__global__ void k_csr_scatter_f32(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                                  const float* __restrict__ v, float* __restrict__ D) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  for (int p = rp[i]; p < rp[i + 1]; ++p) D[(size_t)ci[p] * m + i] = v[p];
}

void mixed_to_dense(const DMat& M, void* dense_dev, hipStream_t st) {
  if (M.kind == K_DENSE) {
    FH_HIP(hipMemcpyAsync(dense_dev, M.d.p, dtype_size(M.dtype) * (size_t)M.m * M.n, hipMemcpyDeviceToDevice, st));
    return;
  }
  FH_CHECK(M.dtype == FH_F32 || M.kind == K_BSR, "to_dense: BF16 CSR not supported");
  FH_HIP(hipMemsetAsync(dense_dev, 0, sizeof(float) * (size_t)M.m * M.n, st));
  if (M.kind == K_BSR) {
    if (M.m > 0)
      hipLaunchKernelGGL(k_bsr_to_dense, dim3(M.m / 32), dim3(256), 0, st, M.m / 32, M.brp.as<int>(),
                         M.bci.as<int>(), M.bdat.as<u16>(), M.split ? 1 : 0, M.m, (float*)dense_dev);
  } else if (M.nnz > 0) {
    hipLaunchKernelGGL(k_csr_scatter_f32, dim3((M.m + 255) / 256), dim3(256), 0, st, M.m, M.rp.as<int>(),
                       M.ci.as<int>(), M.v.as<float>(), (float*)dense_dev);
  }
  FH_LAUNCH_CHECK();
}

// Single product with column-major X and Y (fh_mat_mul_dense).
void mixed_mul_dense(const DMat& S, char op, const void* X, int ldx, int ncols, void* Y, int ldy, const void* alpha,
                     const void* beta, hipStream_t st) {
  const float a = alpha ? *(const float*)alpha : 1.0f, b = beta ? *(const float*)beta : 0.0f;
  if (S.kind == K_BSR) {
    bsr_mul(S, op, X, ldx, ncols, Y, ldy, a, b, st);
    return;
  }
  FH_CHECK(S.kind == K_CSR && S.dtype == FH_F32, "FP32/BF16 products: BSR or FP32 CSR factors only");
  const int in_rows = op == 'N' ? S.n : S.m;
  const int ldr = (ncols + 3) & ~3;
  DBuf R(sizeof(float) * std::max<size_t>((size_t)in_rows * ldr, 1));
  col2row(in_rows, ncols, (const float*)X, ldx, R.as<float>(), ldr, st);
  csr_f32_mul_rm(S, op, R.as<float>(), ldr, ncols, (float*)Y, ldy, true, a, b, st);
  FH_HIP(hipStreamSynchronize(st));  // R returns to the pool
}

// The one-launch BSR chain is checked on its first use in the process, and
// after every use with FH_BSR_CHAIN_CHECK=1
// (then a bad report is an error); a variant whose check fails is not used
// again.  The check synchronizes, so it is skipped while the stream is being
// captured into a graph.
static int g_chain_state[1] = {0};  // 0 unchecked, 1 verified, -1 disabled

static bool chain_usable(int sync) { return g_chain_state[sync] >= 0; }

static bool chain_verified(int sync, int* ctl, hipStream_t st) {
  const char* e = getenv("FH_BSR_CHAIN_CHECK");
  const bool always = e && e[0] == '1';
  if (g_chain_state[sync] == 1 && !always) return true;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  FH_HIP(hipStreamIsCapturing(st, &cs));
  if (cs != hipStreamCaptureStatusNone) return true;
  const int code = bsr_chain_report(ctl, st);
  if (code != 0) {
    if (always)
      throw Error(FH_ERR_NUMERIC, "one-launch BSR chain reported code " + std::to_string(code) +
                                      " (1: a wait ran out, 2: an XCD team did not run)");
    g_chain_state[sync] = -1;
    return false;
  }
  g_chain_state[sync] = 1;
  return true;
}

// Bytes of each intermediate of a BSR or FP32 CSR chain product (*nbuf of
// them): BSR intermediates are tiled, p padded to whole 32-column groups;
// FP32 CSR ones are row-major with ld = p rounded up to 4.
size_t faust_ws_need_mixed(const std::vector<DMatP>& f, char op, int p, int* nbuf, bool act_f32) {
  const int J = (int)f.size();
  const bool bsr = f[0]->kind == K_BSR;
  if (act_f32) {
    for (auto& x : f) FH_CHECK(x->kind == K_BSR && x->dtype == FH_F32, "FP32 activations need FP32 BSR factors");
    int maxr = 0;
    for (auto& x : f) maxr = std::max(maxr, op == 'N' ? x->m : x->n);
    *nbuf = std::min(2, J - 1);
    return sizeof(float) * std::max<size_t>((size_t)maxr * p, 1);  // column-major, ld = rows
  }
  for (auto& x : f) {
    FH_CHECK((x->kind == K_BSR) == bsr, "Faust::multiply: BSR and CSR factors cannot be mixed in one FP32/BF16 chain");
    if (!bsr) FH_CHECK(x->kind == K_CSR && x->dtype == FH_F32, "Faust::multiply: FP32 chains need FP32 CSR factors");
  }
  int maxr = 0;
  for (auto& x : f) maxr = std::max(maxr, op == 'N' ? x->m : x->n);
  maxr = std::max(maxr, (op == 'N') ? f.back()->n : f.front()->m);
  *nbuf = bsr ? std::min(2, J - 1) : std::min(2, J);
  const int ldr = (p + 3) & ~3;
  return bsr ? sizeof(u16) * std::max<size_t>((size_t)maxr * (((size_t)p + 31) / 32 * 32), 1)
             : sizeof(float) * std::max<size_t>((size_t)maxr * ldr, 1);
}

// Chain product for BSR (BF16 activations) or FP32 CSR chains, intermediates
// in ws (faust_ws_need_mixed bytes each).  Returns the number of product
// kernels launched.
static int multiply_mixed_run(const std::vector<DMatP>& f, char op, const void* X, int p, void* Y, hipStream_t st,
                              const ChainWs& ws, FaustTimers* tm, bool act_f32) {
  const int J = (int)f.size();
  const bool bsr = f[0]->kind == K_BSR;
  const int ldr = bsr ? 0 : (p + 3) & ~3;
  int in_rows = (op == 'N') ? f.back()->n : f.front()->m;
  if (act_f32) {
    // FP32 BSR factors with FP32 activations: one k_bsr_f32 launch per factor,
    // column-major FP32 intermediates
    const float* in = static_cast<const float*>(X);
    for (int i = 0; i < J; ++i) {
      const DMat& S = (op == 'N') ? *f[J - 1 - i] : *f[i];
      const int out_rows = (op == 'N') ? S.m : S.n;
      FH_CHECK(((op == 'N') ? S.n : S.m) == in_rows, "Faust::multiply: factor dimensions do not chain");
      float* out = (i == J - 1) ? static_cast<float*>(Y) : static_cast<float*>(ws.p[i & 1]);
      const size_t slot = (op == 'N') ? (size_t)(J - 1 - i) : (size_t)i;
      if (tm) tm->begin(slot, st);
      bsr_mul_f32(S, op, in, in_rows, p, out, out_rows, 1.0f, 0.0f, st);
      if (tm) tm->end(slot, st);
      in = out;
      in_rows = out_rows;
    }
    return J;
  }
  if (bsr && ws.ctl) {
    const int sync = bsr_chain_sync_mode();
    if (sync >= 0 && bsr_chain_fits(f, op) && chain_usable(sync)) {
      if (tm) tm->begin(0, st);
      bsr_chain_launch(f, op, X, p, Y, ws.p[0], ws.p[1], ws.ctl, st);
      if (tm) tm->end(0, st);
      if (chain_verified(sync, ws.ctl, st)) return 1;
      // the chain reported an invalid output: recompute factor by factor
    }
  }
  if (bsr) {
    const void* in = X;
    for (int i = 0; i < J; ++i) {
      const DMat& S = (op == 'N') ? *f[J - 1 - i] : *f[i];
      const int out_rows = (op == 'N') ? S.m : S.n;
      FH_CHECK(((op == 'N') ? S.n : S.m) == in_rows, "Faust::multiply: factor dimensions do not chain");
      void* out = (i == J - 1) ? Y : ws.p[i & 1];
      const size_t slot = (op == 'N') ? (size_t)(J - 1 - i) : (size_t)i;
      if (tm) tm->begin(slot, st);
      // intermediates stay in the tiled layout; X and Y are column-major
      bsr_mul(S, op, in, in_rows, p, out, out_rows, 1.0f, 0.0f, st, i > 0, i < J - 1);
      if (tm) tm->end(slot, st);
      in = out;
      in_rows = out_rows;
    }
    return J;
  }
  // FP32 CSR: transpose the input once, chain in row-major, last factor writes column-major
  if (tm) tm->begin(J, st);
  col2row(in_rows, p, (const float*)X, in_rows, static_cast<float*>(ws.p[0]), ldr, st);
  if (tm) tm->end(J, st);
  const float* in = static_cast<const float*>(ws.p[0]);
  for (int i = 0; i < J; ++i) {
    const DMat& S = (op == 'N') ? *f[J - 1 - i] : *f[i];
    const int out_rows = (op == 'N') ? S.m : S.n;
    FH_CHECK(((op == 'N') ? S.n : S.m) == in_rows, "Faust::multiply: factor dimensions do not chain");
    const bool last = (i == J - 1);
    float* out = last ? (float*)Y : static_cast<float*>(ws.p[(i + 1) & 1]);
    const size_t slot = (op == 'N') ? (size_t)(J - 1 - i) : (size_t)i;
    if (tm) tm->begin(slot, st);
    csr_f32_mul_rm(S, op, in, ldr, p, out, last ? out_rows : ldr, last, 1.0f, 0.0f, st);
    if (tm) tm->end(slot, st);
    in = out;
    in_rows = out_rows;
  }
  return J + 1;
}

// Autotuning of BSR (BF16 activations) and FP32 CSR chains: in AUTO mode, the
// first product of each chain signature (direction, columns, every op'd
// factor's kind, shape, entries and ELL width, FP32 split, device) outside
// graph capture times the execution forms that give the same bits -- BSR
// kernels 5 (shared-block, the default), 1, 2, 4 (ELL, 1 / 2 / 4 column
// tiles per wave) and 0 (row pointers), FP32 CSR one or two rows per thread
// -- each once untimed, then the minimum of three timings of five
// back-to-back products (HIP events on the launch stream), and keeps the
// quickest: the default unless another is at least 3% quicker.  The K-split
// BSR kernel (other bits) and the one-launch BSR chain (opt-in) are not
// candidates.  Environment overrides of the
// kernels (FH_BSR_KERNEL, FH_BSR_XSPLIT, FH_BSR_NT, FH_CSR_RPT, FH_CSR_ELL,
// FH_BSR_CHAIN) turn it off, and so does FH_AUTOTUNE=0.
static bool mixed_env_override() {
  for (const char* v : {"FH_BSR_KERNEL", "FH_BSR_XSPLIT", "FH_BSR_NT", "FH_CSR_RPT", "FH_CSR_ELL", "FH_BSR_CHAIN"})
    if (getenv(v)) return true;
  return false;
}

static std::string mixed_sig(const std::vector<DMatP>& f, char op, int p) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::string k = "mixed:" + std::to_string(dev) + op + ":" + std::to_string(p);
  const bool tr = op != 'N';
  for (const auto& x : f) {
    const int w = x->kind == K_BSR ? (tr ? x->tell_w : x->ell_w) : (tr ? x->csr_tell_w : x->csr_ell_w);
    k += "," + std::to_string(x->kind) + "/" + std::to_string(x->m) + "x" + std::to_string(x->n) + "/" +
         std::to_string(x->kind == K_BSR ? x->bnnz : x->nnz) + "/" + std::to_string(w) + (x->split ? "s" : "");
  }
  return k;
}

int faust_multiply_mixed(const std::vector<DMatP>& f, char op, const void* X, int p, void* Y, hipStream_t st,
                          const ChainWs& ws, FaustTimers* tm, bool act_f32, bool tune) {
  if (act_f32 || mul_mode() != FH_MUL_AUTO || !autotune_on() || mixed_env_override() || f.empty() || p <= 0)
    return multiply_mixed_run(f, op, X, p, Y, st, ws, tm, act_f32);
  const bool bsr = f[0]->kind == K_BSR;
  const std::string key = mixed_sig(f, op, p);
  int mode = FH_MUL_AUTO;
  if (!autotune_get_mode(key, &mode) && tune && !stream_capturing(st)) {
    const std::vector<int> cand =
        bsr ? std::vector<int>{FH_MUL_BSR_KERNEL(5), FH_MUL_BSR_KERNEL(1), FH_MUL_BSR_KERNEL(2),
                               FH_MUL_BSR_KERNEL(4), FH_MUL_BSR_KERNEL(0)}
            : std::vector<int>{FH_MUL_CSR_ROWS1, FH_MUL_CSR_ROWS2};
    hipEvent_t e0, e1;
    FH_HIP(hipEventCreate(&e0));
    FH_HIP(hipEventCreate(&e1));
    float best_ms = 1e30f;
    mode = cand[0];
    // each timing covers MIXED_TUNE_RUNS back-to-back products, the steady
    // state a caller's loop or graph sees (single products timed alone favour
    // kernels that are quicker only from cold caches)
    constexpr int MIXED_TUNE_RUNS = 5;
    for (int c : cand) {
      MulModeScope scope(c);
      multiply_mixed_run(f, op, X, p, Y, st, ws, nullptr, false);
      float m = 1e30f;
      for (int rep = 0; rep < 3; ++rep) {
        FH_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < MIXED_TUNE_RUNS; ++r) multiply_mixed_run(f, op, X, p, Y, st, ws, nullptr, false);
        FH_HIP(hipEventRecord(e1, st));
        FH_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        FH_HIP(hipEventElapsedTime(&ms, e0, e1));
        m = std::min(m, ms);
      }
      if (m < 0.97f * best_ms) {
        best_ms = m;
        mode = c;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    autotune_put_mode(key, mode);
  }
  MulModeScope scope(mode);
  return multiply_mixed_run(f, op, X, p, Y, st, ws, tm, false);
}

// ---------------------------------------------------------------- BSR handle
// Host: blocks arrive column-major (faust_MatBSR.h:120-150); the device keeps
// them in MFMA A-operand order per block, BF16 (or BF16 hi/lo planes for
// FP32), plus the BSR of the transpose for op T/H.
static inline size_t block_slot(int i, int k) {
  return ((size_t)((k >> 4) * 64 + ((k >> 3) & 1) * 32 + i) << 3) + (k & 7);
}
static void put_block(const void* src, int dtype, size_t b, bool transpose, bool split, std::vector<u16>& dst,
                      size_t slot) {
  const size_t planes = split ? 2 : 1;
  u16* out = &dst[slot * 1024 * planes];
  for (int i = 0; i < 32; ++i)
    for (int k = 0; k < 32; ++k) {
      // entry (i, k) of the stored block (of B or of B^T)
      const int r = transpose ? k : i, c = transpose ? i : k;
      const size_t e = b * 1024 + (size_t)c * 32 + r;  // column-major source
      float v;
      if (dtype == FH_BF16)
        v = host_bf2f(static_cast<const u16*>(src)[e]);
      else
        v = static_cast<const float*>(src)[e];
      const u16 hi = host_f2bf(v);
      out[block_slot(i, k)] = hi;
      if (split) out[1024 + block_slot(i, k)] = host_f2bf(v - host_bf2f(hi));
    }
}

// BSR handle back to host (fh_mat_tocpu): block row pointers, block column
// indices and the blocks column-major one after the other (the layout
// fh_mat_bsr_create takes, faust_MatBSR.h:120-150).  Values in the handle's
// type: BF16 bit patterns, or for FP32 handles hi + lo of the stored pair
// (within 2^-16 relative of the uploaded value).  Synchronous.
void bsr_to_host(const DMat& M, void* bdata, int32_t* browptr, int32_t* bcolinds) {
  FH_CHECK(M.kind == K_BSR, "bsr_to_host: not a BSR matrix");
  const int nbr = M.m / 32, bnnz = M.bnnz;
  const size_t planes = M.split ? 2 : 1;
  if (browptr) FH_HIP(hipMemcpy(browptr, M.brp.p, sizeof(int) * (nbr + 1), hipMemcpyDeviceToHost));
  if (bcolinds && bnnz) FH_HIP(hipMemcpy(bcolinds, M.bci.p, sizeof(int) * bnnz, hipMemcpyDeviceToHost));
  if (!bdata || !bnnz) return;
  if (M.dtype == FH_F32 && M.f32dat.p) {  // the exact FP32 blocks (operand order -> column-major)
    std::vector<float> f((size_t)bnnz * 1024);
    FH_HIP(hipMemcpy(f.data(), M.f32dat.p, sizeof(float) * f.size(), hipMemcpyDeviceToHost));
    for (size_t b = 0; b < (size_t)bnnz; ++b)
      for (int l = 0; l < 64; ++l)
        for (int q = 0; q < 16; ++q) {
          const int r = l & 31, c = (l >> 5) * 16 + q;
          static_cast<float*>(bdata)[b * 1024 + (size_t)c * 32 + r] = f[b * 1024 + l * 16 + q];
        }
    return;
  }
  std::vector<u16> dat((size_t)bnnz * 1024 * planes);
  FH_HIP(hipMemcpy(dat.data(), M.bdat.p, sizeof(u16) * dat.size(), hipMemcpyDeviceToHost));
  for (size_t b = 0; b < (size_t)bnnz; ++b) {
    const u16* blk = &dat[b * 1024 * planes];
    for (int k = 0; k < 32; ++k)
      for (int i = 0; i < 32; ++i) {
        const size_t e = b * 1024 + (size_t)k * 32 + i;  // column-major destination
        const size_t s = block_slot(i, k);
        if (M.dtype == FH_BF16)
          static_cast<u16*>(bdata)[e] = blk[s];
        else
          static_cast<float*>(bdata)[e] = host_bf2f(blk[s]) + host_bf2f(blk[1024 + s]);
      }
  }
}

// ELL slots of a block-row structure: width = widest block row rounded up to
// a multiple of 4 (0 when above BSR_ELL_MAX); slot (br, q) = (rp[br] + q,
// ci[rp[br] + q]), or (zero_block, 0) past the end of the row, zero_block
// being an all-zero block stored after the nnz blocks.
static int build_ell(int nrows, const int32_t* rp, const int32_t* ci, int zero_block, std::vector<int32_t>& ell) {
  int w = 0;
  for (int r = 0; r < nrows; ++r) w = std::max(w, rp[r + 1] - rp[r]);
  if (w > BSR_ELL_MAX) return 0;
  w = std::max(4, (w + 3) / 4 * 4);
  ell.assign((size_t)nrows * w * 2, 0);
  for (int r = 0; r < nrows; ++r)
    for (int q = 0; q < w; ++q) {
      const int b = rp[r] + q;
      const bool on = b < rp[r + 1];
      ell[((size_t)r * w + q) * 2] = on ? b : zero_block;
      ell[((size_t)r * w + q) * 2 + 1] = on ? ci[b] : 0;
    }
  return w;
}

DMatP make_bsr(int dtype, int m, int n, int bnnz, const int32_t* browptr, const int32_t* bcolinds,
               const void* bdata, hipStream_t st) {
  FH_CHECK(dtype == FH_BF16 || dtype == FH_F32, "BSR factors are BF16 or FP32");
  FH_CHECK(m % 32 == 0 && n % 32 == 0, "BSR: dimensions must be multiples of the block size");
  const int nbr = m / 32, nbc = n / 32;
  FH_CHECK(browptr[0] == 0 && browptr[nbr] == bnnz, "BSR: browptr inconsistent with bnnz");
  for (int br = 0; br < nbr; ++br) {
    FH_CHECK(browptr[br] <= browptr[br + 1], "BSR: browptr not non-decreasing");
    for (int b = browptr[br]; b < browptr[br + 1]; ++b)
      FH_CHECK(bcolinds[b] >= 0 && bcolinds[b] < nbc, "BSR: block column index out of range");
  }
  const bool split = dtype == FH_F32;
  const size_t planes = split ? 2 : 1;
  auto M = std::make_shared<DMat>();
  M->kind = K_BSR;
  M->dtype = dtype;
  M->m = m;
  M->n = n;
  M->bm = M->bn = 32;
  M->bnnz = bnnz;
  M->split = split;
  M->nnz = (int64_t)bnnz * 1024;
  // bnnz blocks plus one all-zero block (the target of empty ELL slots)
  std::vector<u16> dat((size_t)(bnnz + 1) * 1024 * planes), tdat(dat.size());
  for (int b = 0; b < bnnz; ++b) put_block(bdata, dtype, b, false, split, dat, b);
  // transposed BSR: block (br, bc) -> (bc, br) with the block transposed
  std::vector<int32_t> tbrp(nbc + 1, 0), tbci(std::max(bnnz, 1));
  for (int b = 0; b < bnnz; ++b) tbrp[bcolinds[b] + 1]++;
  for (int c = 0; c < nbc; ++c) tbrp[c + 1] += tbrp[c];
  std::vector<int32_t> fill(tbrp.begin(), tbrp.end() - 1);
  for (int br = 0; br < nbr; ++br)
    for (int b = browptr[br]; b < browptr[br + 1]; ++b) {
      const int slot = fill[bcolinds[b]]++;
      tbci[slot] = br;
      put_block(bdata, dtype, b, true, split, tdat, slot);
    }
  auto up = [&](DBuf& d, const void* h, size_t bytes) {
    d.alloc(std::max<size_t>(bytes, 4));
    if (bytes) FH_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, st));
  };
  std::vector<float> f32, tf32;
  if (dtype == FH_F32) {
    // exact FP32 blocks for the FP32-activation product, same block order
    f32.assign((size_t)(bnnz + 1) * 1024, 0.0f);
    tf32.assign(f32.size(), 0.0f);
    std::vector<int> tslot_of(std::max(bnnz, 1));
    {
      std::vector<int32_t> tf(tbrp.begin(), tbrp.end() - 1);
      for (int br = 0; br < nbr; ++br)
        for (int b = browptr[br]; b < browptr[br + 1]; ++b) tslot_of[b] = tf[bcolinds[b]]++;
    }
    const float* src = static_cast<const float*>(bdata);
    for (int b = 0; b < bnnz; ++b)
      for (int l = 0; l < 64; ++l)
        for (int q = 0; q < 16; ++q) {
          const int r = l & 31, c = (l >> 5) * 16 + q;
          f32[(size_t)b * 1024 + l * 16 + q] = src[(size_t)b * 1024 + (size_t)c * 32 + r];          // B(r, c)
          tf32[(size_t)tslot_of[b] * 1024 + l * 16 + q] = src[(size_t)b * 1024 + (size_t)r * 32 + c];  // B(c, r)
        }
  }
  up(M->brp, browptr, sizeof(int) * (nbr + 1));
  if (dtype == FH_F32) {
    up(M->f32dat, f32.data(), sizeof(float) * f32.size());
    up(M->tf32dat, tf32.data(), sizeof(float) * tf32.size());
  }
  up(M->bci, bcolinds, sizeof(int) * bnnz);
  up(M->bdat, dat.data(), sizeof(u16) * (size_t)(bnnz + 1) * 1024 * planes);
  up(M->tbrp, tbrp.data(), sizeof(int) * (nbc + 1));
  up(M->tbci, tbci.data(), sizeof(int) * bnnz);
  up(M->tbdat, tdat.data(), sizeof(u16) * (size_t)(bnnz + 1) * 1024 * planes);
  std::vector<int32_t> ell, tell;
  M->ell_w = build_ell(nbr, browptr, bcolinds, bnnz, ell);
  M->tell_w = build_ell(nbc, tbrp.data(), tbci.data(), bnnz, tell);
  if (M->ell_w) up(M->ell, ell.data(), sizeof(int32_t) * ell.size());
  if (M->tell_w) up(M->tell, tell.data(), sizeof(int32_t) * tell.size());
  // blocks in ELL order (a slot's block index is ell[2 * slot])
  const size_t blk_elems = (size_t)1024 * planes;
  auto ell_blocks = [&](int nrows, int w, const std::vector<int32_t>& e, const std::vector<u16>& src,
                        std::vector<u16>& out) {
    out.assign((size_t)nrows * w * blk_elems, 0);
    for (size_t slot = 0; slot < (size_t)nrows * w; ++slot)
      std::copy(src.begin() + (size_t)e[2 * slot] * blk_elems, src.begin() + (size_t)(e[2 * slot] + 1) * blk_elems,
                out.begin() + slot * blk_elems);
  };
  std::vector<u16> edat;
  if (M->ell_w) {
    ell_blocks(nbr, M->ell_w, ell, dat, edat);
    up(M->ebdat, edat.data(), sizeof(u16) * edat.size());
  }
  if (M->tell_w) {
    std::vector<u16> tedat;
    ell_blocks(nbc, M->tell_w, tell, tdat, tedat);
    up(M->tebdat, tedat.data(), sizeof(u16) * tedat.size());
    FH_HIP(hipStreamSynchronize(st));  // tedat goes out of scope
  }
  FH_HIP(hipStreamSynchronize(st));
  return M;
}

// BSR with other block sizes: expanded to an FP32 CSR factor.
static DMatP bsr_as_csr(int dtype, int m, int n, int bm, int bn, int bnnz, const int32_t* browptr,
                        const int32_t* bcolinds, const void* bdata, hipStream_t st) {
  FH_CHECK(bm > 0 && bn > 0 && m % bm == 0 && n % bn == 0, "BSR: dimensions must be multiples of the block size");
  const int nbr = m / bm;
  std::vector<int32_t> rp(m + 1, 0), ci;
  std::vector<float> v;
  std::vector<double> v64;  // FP64 blocks keep their values (and the canonical FP64 products)
  for (int br = 0; br < nbr; ++br) {
    std::vector<int> order;
    for (int b = browptr[br]; b < browptr[br + 1]; ++b) order.push_back(b);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return bcolinds[x] < bcolinds[y]; });
    for (int i = 0; i < bm; ++i) {
      for (int b : order)
        for (int k = 0; k < bn; ++k) {
          const size_t e = (size_t)b * bm * bn + (size_t)k * bm + i;
          ci.push_back(bcolinds[b] * bn + k);
          if (dtype == FH_F64)
            v64.push_back(static_cast<const double*>(bdata)[e]);
          else
            v.push_back(dtype == FH_BF16 ? host_bf2f(static_cast<const u16*>(bdata)[e])
                                         : static_cast<const float*>(bdata)[e]);
        }
      rp[br * bm + i + 1] = (int32_t)ci.size();
    }
  }
  if (dtype == FH_F64)
    return make_csr_from_host(FH_F64, m, n, (int)ci.size(), rp.data(), ci.data(), v64.data(), st);
  return make_csr_from_host(FH_F32, m, n, (int)ci.size(), rp.data(), ci.data(), v.data(), st);
}

}  // namespace fh

using namespace fh;

extern "C" int fh_mat_bsr_create(int dtype, int nrows, int ncols, int bm, int bn, int bnnz, const int32_t* browptr,
                                 const int32_t* bcolinds, const void* bdata, int src_on_device, fh_mat_t* out) {
  FH_API_BEGIN
  FH_CHECK(out && browptr && (bnnz == 0 || (bcolinds && bdata)), "null argument");
  FH_CHECK(bm > 0 && bn > 0 && nrows % bm == 0 && ncols % bn == 0, "BSR: bad block size");
  const size_t esz = dtype_size(dtype);
  std::vector<int32_t> hp, hc;
  std::vector<char> hd;
  if (src_on_device) {
    const int nbr = nrows / bm;
    hp.resize(nbr + 1);
    hc.resize(bnnz);
    hd.resize(esz * (size_t)bnnz * bm * bn);
    FH_HIP(hipMemcpy(hp.data(), browptr, sizeof(int) * (nbr + 1), hipMemcpyDeviceToHost));
    if (bnnz) {
      FH_HIP(hipMemcpy(hc.data(), bcolinds, sizeof(int) * bnnz, hipMemcpyDeviceToHost));
      FH_HIP(hipMemcpy(hd.data(), bdata, hd.size(), hipMemcpyDeviceToHost));
    }
    browptr = hp.data();
    bcolinds = hc.data();
    bdata = hd.data();
  }
  if (bm == 32 && bn == 32 && dtype != FH_F64)  // FP64 BSR factors run as CSR (canonical FP64 order)
    *out = new fh_mat_s{make_bsr(dtype, nrows, ncols, bnnz, browptr, bcolinds, bdata, nullptr)};
  else
    *out = new fh_mat_s{bsr_as_csr(dtype, nrows, ncols, bm, bn, bnnz, browptr, bcolinds, bdata, nullptr)};
  FH_API_END
}
