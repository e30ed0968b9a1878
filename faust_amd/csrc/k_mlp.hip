// k_mlp.hip -- the warm-start MLP (in -> h1 -> h2 -> out, GELU) trained and
// run with the project's own kernels: BF16 operands on MFMA
// (v_mfma_f32_32x32x16_bf16) with FP32 accumulation, FP32 master weights,
// FP32 activations and gradients, Adam.
//
// One GEMM kernel serves every product: C[M x N] (FP32) = op(A) op(B) with
// BF16 operands addressed by (row, col) strides, staged through LDS as
// k-contiguous tiles so that each MFMA operand is one 16-byte LDS read.  A
// workgroup of 4 waves computes a 64 x 64 tile (one 32 x 32 accumulator per
// wave); split-K across workgroups adds partial tiles with FP32 atomics for
// the very long reductions (65536-wide inputs).
#include <cmath>
#include <random>
#include <vector>

#include "fh_api_common.hpp"

namespace fh {

typedef __bf16 mbf16x8 __attribute__((ext_vector_type(8)));
typedef float mf32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16;

__device__ __forceinline__ u16 mlp_f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (u16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}

// element (i, k) of A at A[i*sAi + k*sAk]; element (k, j) of B at B[k*sBk + j*sBj]
__global__ __launch_bounds__(256) void k_mlp_gemm(int M, int N, int K, const u16* __restrict__ A, long sAi, long sAk,
                                                  const u16* __restrict__ B, long sBk, long sBj,
                                                  float* __restrict__ C, long ldc, int ksplit, int accumulate) {
  __shared__ u16 As[64][16 + 8];  // [row][k]
  __shared__ u16 Bs[64][16 + 8];  // [col][k]
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int ntn = (N + 63) / 64;
  const int tile = blockIdx.x;
  const int i0 = (tile / ntn) * 64, j0 = (tile % ntn) * 64;
  const int kchunk = ((K + ksplit - 1) / ksplit + 15) & ~15;
  const int kb = blockIdx.y * kchunk, ke = min(K, kb + kchunk);
  const int wy = wave >> 1, wx = wave & 1;
  mf32x16 acc = {};
  for (int k0 = kb; k0 < ke; k0 += 16) {
    // 64 x 16 tiles, 4 elements per thread
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = t + 256 * q;
      const int r = e >> 4, kk = e & 15;
      const int gi = i0 + r, gk = k0 + kk;
      As[r][kk] = (gi < M && gk < ke) ? A[(long)gi * sAi + (long)gk * sAk] : (u16)0;
      const int gj = j0 + r;
      Bs[r][kk] = (gj < N && gk < ke) ? B[(long)gk * sBk + (long)gj * sBj] : (u16)0;
    }
    __syncthreads();
    const mbf16x8 a = *reinterpret_cast<const mbf16x8*>(&As[wy * 32 + (lane & 31)][8 * (lane >> 5)]);
    const mbf16x8 b = *reinterpret_cast<const mbf16x8*>(&Bs[wx * 32 + (lane & 31)][8 * (lane >> 5)]);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    __syncthreads();
  }
  const int j = j0 + wx * 32 + (lane & 31);
  if (j >= N) return;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = i0 + wy * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (i < M) {
      float* c = C + (long)i * ldc + j;
      if (ksplit > 1 || accumulate)
        atomicAdd(c, acc[r]);
      else
        *c = acc[r];
    }
  }
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

// Z[b, j] += bias[j]; H = act(Z) in FP32 and BF16 (act 0: identity, 1: GELU)
__global__ void k_bias_act(float* __restrict__ Z, const float* __restrict__ bias, int B, int N, int act,
                           float* __restrict__ H, u16* __restrict__ Hb) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)B * N) return;
  const float z = Z[i] + bias[i % N];
  Z[i] = z;
  const float h = act ? gelu(z) : z;
  if (H) H[i] = h;
  if (Hb) Hb[i] = mlp_f2bf(h);
}

__global__ void k_f32_to_bf16(const float* __restrict__ x, u16* __restrict__ y, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = mlp_f2bf(x[i]);
}

// dY = 2/(B N) (Y - T); loss partial sums
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, long n, float scale,
                      float* __restrict__ dY, u16* __restrict__ dYb, float* __restrict__ loss) {
  __shared__ float red[256];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float d = 0.f;
  if (i < n) {
    d = Y[i] - T[i];
    const float g = 2.0f * scale * d;
    dY[i] = g;
    dYb[i] = mlp_f2bf(g);
  }
  red[threadIdx.x] = d * d;
  __syncthreads();
  for (int s = 128; s >= 1; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] * scale);
}

// dZ = dH * gelu'(Z) (in place on dH), BF16 copy
__global__ void k_gelu_back(float* __restrict__ dH, const float* __restrict__ Z, long n, u16* __restrict__ dZb) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = dH[i] * gelu_grad(Z[i]);
  dH[i] = g;
  dZb[i] = mlp_f2bf(g);
}

// db[j] = sum_b dZ[b, j]
__global__ void k_bias_grad(const float* __restrict__ dZ, int B, int N, float* __restrict__ db) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  float s = 0.f;
  for (int b = 0; b < B; ++b) s += dZ[(long)b * N + j];
  db[j] = s;
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2,
                       u16* __restrict__ wb) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = mlp_f2bf(wi);
}

static inline dim3 g1d(long n) { return dim3((unsigned)((n + 255) / 256)); }

// C = op(A) op(B) (+C when accumulate); picks a split-K factor for long K
static void mlp_gemm(int M, int N, int K, const u16* A, long sAi, long sAk, const u16* B, long sBk, long sBj,
                     float* C, long ldc, bool accumulate, hipStream_t st) {
  const long tiles = (long)((M + 63) / 64) * ((N + 63) / 64);
  int ksplit = 1;
  while (tiles * ksplit < 1024 && K / (ksplit * 2) >= 256) ksplit *= 2;
  if (ksplit > 1 && !accumulate) FH_HIP(hipMemsetAsync(C, 0, sizeof(float) * (size_t)M * ldc, st));
  hipLaunchKernelGGL(k_mlp_gemm, dim3((unsigned)tiles, ksplit), dim3(256), 0, st, M, N, K, A, sAi, sAk, B, sBk,
                     sBj, C, ldc, ksplit, accumulate ? 1 : 0);
  FH_LAUNCH_CHECK();
}

}  // namespace fh

using namespace fh;

// Layer l: W_l (in_l x out_l, row-major), b_l (out_l).
struct fh_mlp_s {
  int dims[4];
  std::vector<DBuf> W, Wb, b, gW, gb, mW, vW, mb, vb;  // per layer
  DBuf Z[3], H[3], Hb[3], dH[3], dZb[3], Xb, loss;
  int cap_batch = 0;
  long step = 0;
};

static void mlp_workspace(fh_mlp_s& M, int B) {
  if (B <= M.cap_batch) return;
  for (int l = 0; l < 3; ++l) {
    const size_t n = (size_t)B * M.dims[l + 1];
    M.Z[l].alloc(sizeof(float) * n);
    M.H[l].alloc(sizeof(float) * n);
    M.Hb[l].alloc(sizeof(u16) * n);
    M.dH[l].alloc(sizeof(float) * n);
    M.dZb[l].alloc(sizeof(u16) * n);
  }
  M.Xb.alloc(sizeof(u16) * (size_t)B * M.dims[0]);
  M.loss.alloc(sizeof(float));
  M.cap_batch = B;
}

static void mlp_forward(fh_mlp_s& M, const float* X, int B, hipStream_t st) {
  mlp_workspace(M, B);
  const long nx = (long)B * M.dims[0];
  hipLaunchKernelGGL(k_f32_to_bf16, g1d(nx), dim3(256), 0, st, X, M.Xb.as<u16>(), nx);
  FH_LAUNCH_CHECK();
  const u16* in = M.Xb.as<u16>();
  for (int l = 0; l < 3; ++l) {
    const int K = M.dims[l], N = M.dims[l + 1];
    // Z = in (B x K, row-major) * W (K x N, row-major)
    mlp_gemm(B, N, K, in, K, 1, M.Wb[l].as<u16>(), N, 1, M.Z[l].as<float>(), N, false, st);
    hipLaunchKernelGGL(k_bias_act, g1d((long)B * N), dim3(256), 0, st, M.Z[l].as<float>(), M.b[l].as<float>(), B, N,
                       l < 2 ? 1 : 0, M.H[l].as<float>(), M.Hb[l].as<u16>());
    FH_LAUNCH_CHECK();
    in = M.Hb[l].as<u16>();
  }
}

// gradients of the MSE loss into gW/gb; returns the loss (synchronizes)
static float mlp_backward(fh_mlp_s& M, const float* T, int B, hipStream_t st) {
  FH_HIP(hipMemsetAsync(M.loss.p, 0, sizeof(float), st));
  const long ny = (long)B * M.dims[3];
  hipLaunchKernelGGL(k_mse, g1d(ny), dim3(256), 0, st, M.H[2].as<float>(), T, ny, 1.0f / (float)ny,
                     M.dH[2].as<float>(), M.dZb[2].as<u16>(), M.loss.as<float>());
  FH_LAUNCH_CHECK();
  for (int l = 2; l >= 0; --l) {
    const int K = M.dims[l], N = M.dims[l + 1];
    if (l < 2) {
      hipLaunchKernelGGL(k_gelu_back, g1d((long)B * N), dim3(256), 0, st, M.dH[l].as<float>(), M.Z[l].as<float>(),
                         (long)B * N, M.dZb[l].as<u16>());
      FH_LAUNCH_CHECK();
    }
    const u16* in = l == 0 ? M.Xb.as<u16>() : M.Hb[l - 1].as<u16>();
    // gW = in^T (K x B) * dZ (B x N)
    mlp_gemm(K, N, B, in, 1, K, M.dZb[l].as<u16>(), N, 1, M.gW[l].as<float>(), N, false, st);
    hipLaunchKernelGGL(k_bias_grad, g1d(N), dim3(256), 0, st, M.dH[l].as<float>(), B, N, M.gb[l].as<float>());
    FH_LAUNCH_CHECK();
    if (l > 0) {
      // dH_{l-1} = dZ (B x N) * W^T (N x K)
      mlp_gemm(B, K, N, M.dZb[l].as<u16>(), N, 1, M.Wb[l].as<u16>(), 1, N, M.dH[l - 1].as<float>(), K, false, st);
    }
  }
  float h = 0.f;
  FH_HIP(hipMemcpyAsync(&h, M.loss.p, sizeof(float), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  return h;
}

static void mlp_upload_bf16(fh_mlp_s& M, int l, hipStream_t st) {
  const long n = (long)M.dims[l] * M.dims[l + 1];
  hipLaunchKernelGGL(k_f32_to_bf16, g1d(n), dim3(256), 0, st, M.W[l].as<float>(), M.Wb[l].as<u16>(), n);
  FH_LAUNCH_CHECK();
}

extern "C" {

int fh_mlp_create(int n_in, int n_h1, int n_h2, int n_out, uint64_t seed, fh_mlp_t* out) {
  FH_API_BEGIN
  FH_CHECK(out && n_in > 0 && n_h1 > 0 && n_h2 > 0 && n_out > 0, "bad argument");
  auto M = std::make_unique<fh_mlp_s>();
  M->dims[0] = n_in;
  M->dims[1] = n_h1;
  M->dims[2] = n_h2;
  M->dims[3] = n_out;
  std::mt19937_64 rng(seed);
  for (auto* v : {&M->W, &M->Wb, &M->b, &M->gW, &M->gb, &M->mW, &M->vW, &M->mb, &M->vb}) v->resize(3);
  for (int l = 0; l < 3; ++l) {
    const long K = M->dims[l], N = M->dims[l + 1];
    // He initialisation for GELU layers, Glorot-scale for the linear output
    const double sd = l < 2 ? std::sqrt(2.0 / K) : std::sqrt(1.0 / K);
    std::normal_distribution<double> nd(0.0, sd);
    std::vector<float> w(K * N);
    for (auto& x : w) x = (float)nd(rng);
    M->W[l].alloc(sizeof(float) * K * N);
    M->Wb[l].alloc(sizeof(u16) * K * N);
    M->gW[l].alloc(sizeof(float) * K * N);
    M->mW[l].alloc(sizeof(float) * K * N);
    M->vW[l].alloc(sizeof(float) * K * N);
    M->b[l].alloc(sizeof(float) * N);
    M->gb[l].alloc(sizeof(float) * N);
    M->mb[l].alloc(sizeof(float) * N);
    M->vb[l].alloc(sizeof(float) * N);
    FH_HIP(hipMemcpy(M->W[l].p, w.data(), sizeof(float) * K * N, hipMemcpyHostToDevice));
    FH_HIP(hipMemset(M->b[l].p, 0, sizeof(float) * N));
    FH_HIP(hipMemset(M->mW[l].p, 0, sizeof(float) * K * N));
    FH_HIP(hipMemset(M->vW[l].p, 0, sizeof(float) * K * N));
    FH_HIP(hipMemset(M->mb[l].p, 0, sizeof(float) * N));
    FH_HIP(hipMemset(M->vb[l].p, 0, sizeof(float) * N));
    mlp_upload_bf16(*M, l, nullptr);
  }
  FH_HIP(hipDeviceSynchronize());
  *out = M.release();
  FH_API_END
}

int fh_mlp_free(fh_mlp_t M) {
  FH_API_BEGIN
  delete M;
  FH_API_END
}

// Y (B x n_out, row-major FP32, device) = MLP(X (B x n_in row-major FP32, device))
int fh_mlp_forward(fh_mlp_t M, const float* X, int B, float* Y, void* stream) {
  FH_API_BEGIN
  FH_CHECK(M && X && Y && B > 0, "bad argument");
  hipStream_t st = (hipStream_t)stream;
  mlp_forward(*M, X, B, st);
  FH_HIP(hipMemcpyAsync(Y, M->H[2].p, sizeof(float) * (size_t)B * M->dims[3], hipMemcpyDeviceToDevice, st));
  FH_API_END
}

// Gradients of the mean squared error (dims: all parameters) for one batch;
// `loss` receives the loss.  Synchronizes.
int fh_mlp_gradients(fh_mlp_t M, const float* X, const float* T, int B, float* loss, void* stream) {
  FH_API_BEGIN
  FH_CHECK(M && X && T && B > 0, "bad argument");
  hipStream_t st = (hipStream_t)stream;
  mlp_forward(*M, X, B, st);
  const float l = mlp_backward(*M, T, B, st);
  if (loss) *loss = l;
  FH_API_END
}

// One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the gradients of one batch.
int fh_mlp_train_step(fh_mlp_t M, const float* X, const float* T, int B, float lr, float* loss, void* stream) {
  FH_API_BEGIN
  FH_CHECK(M && X && T && B > 0, "bad argument");
  hipStream_t st = (hipStream_t)stream;
  mlp_forward(*M, X, B, st);
  const float l = mlp_backward(*M, T, B, st);
  M->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)M->step), c2 = 1.f - std::pow(b2, (float)M->step);
  for (int li = 0; li < 3; ++li) {
    const long nw = (long)M->dims[li] * M->dims[li + 1], nb = M->dims[li + 1];
    hipLaunchKernelGGL(k_adam, g1d(nw), dim3(256), 0, st, M->W[li].as<float>(), M->gW[li].as<float>(),
                       M->mW[li].as<float>(), M->vW[li].as<float>(), nw, lr, b1, b2, 1e-8f, c1, c2,
                       M->Wb[li].as<u16>());
    hipLaunchKernelGGL(k_adam, g1d(nb), dim3(256), 0, st, M->b[li].as<float>(), M->gb[li].as<float>(),
                       M->mb[li].as<float>(), M->vb[li].as<float>(), nb, lr, b1, b2, 1e-8f, c1, c2, (u16*)nullptr);
    FH_LAUNCH_CHECK();
  }
  if (loss) *loss = l;
  FH_API_END
}

// Host copies of layer l's parameters (W: in x out row-major) or gradients.
int fh_mlp_get(fh_mlp_t M, int layer, int what, float* W, float* b) {
  FH_API_BEGIN
  FH_CHECK(M && layer >= 0 && layer < 3, "bad argument");
  const size_t nw = (size_t)M->dims[layer] * M->dims[layer + 1], nb = M->dims[layer + 1];
  const DBuf& w = what == 0 ? M->W[layer] : M->gW[layer];
  const DBuf& bb = what == 0 ? M->b[layer] : M->gb[layer];
  FH_HIP(hipDeviceSynchronize());
  if (W) FH_HIP(hipMemcpy(W, w.p, sizeof(float) * nw, hipMemcpyDeviceToHost));
  if (b) FH_HIP(hipMemcpy(b, bb.p, sizeof(float) * nb, hipMemcpyDeviceToHost));
  FH_API_END
}

int fh_mlp_set(fh_mlp_t M, int layer, const float* W, const float* b) {
  FH_API_BEGIN
  FH_CHECK(M && layer >= 0 && layer < 3, "bad argument");
  const size_t nw = (size_t)M->dims[layer] * M->dims[layer + 1], nb = M->dims[layer + 1];
  if (W) FH_HIP(hipMemcpy(M->W[layer].p, W, sizeof(float) * nw, hipMemcpyHostToDevice));
  if (b) FH_HIP(hipMemcpy(M->b[layer].p, b, sizeof(float) * nb, hipMemcpyHostToDevice));
  mlp_upload_bf16(*M, layer, nullptr);
  FH_HIP(hipDeviceSynchronize());
  FH_API_END
}

}  // extern "C"
