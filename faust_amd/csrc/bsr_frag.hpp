// bsr_frag.hpp -- device helpers shared by the BSR kernels (k_mixed.hip,
// k_bsr_chain.hip): BF16 vector types, conversions, the fragment-tiled
// activation layout and the BF16 epilogue.
#pragma once
#include <hip/hip_runtime.h>

namespace fh {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16;
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float(((unsigned)h) << 16); }
// Two FP32 -> packed BF16 (lo = a, hi = b), round to nearest even: one
// v_cvt_pk_bf16_f32 on gfx950 (inf stays inf, NaN stays NaN).
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned f2bf2(float a, float b) {
  bf16x2 r;
  r[0] = (__bf16)a;
  r[1] = (__bf16)b;
  return __builtin_bit_cast(unsigned, r);
}
// ---------------------------------------------------------------- BSR MFMA
// Activation layouts: column-major (ld), or "tiled" (fragment order): the
// columns are cut into G = ceil(p/32) groups of 32 (the last one padded), and
// the 32 x 32 tile of block-row rb, column group g is 1024 consecutive
// entries at (rb*G + g)*1024, entry (r, c) of the tile at
//   ((r / 16)*64 + ((r / 8) % 2)*32 + c)*8 + r % 8.
// That is the MFMA B-operand order: for the K half h the 64 lanes read 16
// bytes each at (h*64 + lane)*8, one contiguous 1 KiB, and the accumulator of
// a 32x32 output tile leaves as four contiguous 512-byte stores.  Chains keep
// their intermediates tiled; only the first factor reads and the last factor
// writes column-major.  Padding columns (p..32G) are never written and only
// feed accumulator columns that are never stored.
__device__ __forceinline__ size_t frag_tile(int rb, int G, int g) { return ((size_t)rb * G + g) * 1024; }
// offset, inside a tile, of the 4 rows 8*gg + 4*kh .. +3 of column c (one
// accumulator quarter of lane (c, kh))
__device__ __forceinline__ int frag_store_off(int gg, int kh, int c) {
  return (((gg >> 1) * 64 + (gg & 1) * 32 + c) << 3) + 4 * kh;
}
// Wave work: one block-row x 64 columns (two 32x32 accumulators sharing the
// A operand); workgroup: 4 block-rows x 64 columns.
// Epilogue of 4 consecutive rows of one column: y = bf16(alpha*acc) or, with
// beta != 0 (a wave-uniform test), bf16(fma(beta, y_old, alpha*acc)); the
// old values are one 8-byte load.
template <bool NT = false>
__device__ __forceinline__ void bsr_store4(u16* yq, float a0, float a1, float a2, float a3, float alpha,
                                           float beta) {
  float v0 = alpha * a0, v1 = alpha * a1, v2 = alpha * a2, v3 = alpha * a3;
  if (beta != 0.0f) {
    const uint2 old = *reinterpret_cast<const uint2*>(yq);
    v0 = fmaf(beta, __uint_as_float(old.x << 16), v0);
    v1 = fmaf(beta, __uint_as_float(old.x & 0xffff0000u), v1);
    v2 = fmaf(beta, __uint_as_float(old.y << 16), v2);
    v3 = fmaf(beta, __uint_as_float(old.y & 0xffff0000u), v3);
  }
  const unsigned wx = f2bf2(v0, v1), wy = f2bf2(v2, v3);
  if (NT) {
    u32x2 wv = {wx, wy};
    __builtin_nontemporal_store(wv, reinterpret_cast<u32x2*>(yq));
  } else {
    uint2 w;
    w.x = wx;
    w.y = wy;
    *reinterpret_cast<uint2*>(yq) = w;
  }
}

}  // namespace fh
