// k_power.hip -- power iteration on a factor chain in one persistent launch.
//
// The reference (power_iteration, faust_linear_algebra.hpp:808-898) runs up
// to 100 dependent rounds of "normalise, apply the F F^H chain, Rayleigh
// quotient", each a string of tiny mat-vecs.  Here a single 1024-thread
// workgroup keeps both vectors in LDS and runs every round, every mat-vec and
// the stopping test on the device: no launch or host round trip per step.
// Mat-vecs are one fma chain per output row (k ascending); norms and dots use
// the canonical 256-lane chunk reduction, so the result matches the oracle
// bit for bit.
//
// The chains of PALM4MSA's Lipschitz constants are long strings of small
// sparse factors (a global update of a 10-factor Faust applies up to 18
// factors of 2 entries per row each round, for 100 rounds when the estimate
// does not settle), so a round is a string of dependent latencies, not of
// bytes.  The kernel therefore copies the chain's row pointers into LDS once
// per call, reads the operator descriptors through scalar loads, and loads the
// entries of a thread's rows for the operator after next while it computes the
// current one (rows of up to PW_PF entries; longer rows read theirs in the
// loop), with barriers that order only the LDS vectors, so those loads stay in
// flight across them.  Measured on the 1024 Hadamard hierarchy (100
// iterations per PALM4MSA): 0.71 us per operator and 3.3 us of norm,
// normalisation and Rayleigh quotient per round (11.5 s in this kernel); the
// whole hierarchy takes 24.1 s, against 26.8 s when each thread read its row
// pointers and entries from global memory as it needed them.  Chains whose
// operators are all sparse with rows of at most PW_PF entries run a second
// copy of the round loop without the loops over long rows and dense
// operators: with those in the loop, the compiler waited for every load in
// flight at the start of each operator, the prefetched set included
// (hierarchy 24.3 -> 23.6 s).
#include <cstdlib>
#include <map>
#include <mutex>

#include "fh_internal.hpp"

namespace fh {

constexpr int PW_THREADS = 1024;
constexpr int PW_VMAX = 2048;  // max vector length kept in LDS (3 vectors: 48 KiB)

struct PwOp {
  int kind;   // 0 dense, 1 row-compressed sparse
  int trans;  // dense only: op(A) = A^T
  int rows, cols;
  const int* rp;
  const int* ci;
  const double* v;  // sparse values or dense data (ld = rows of A)
  int lda;
  int rpo;      // sparse only: offset of the row pointers' copy in dynamic LDS, or -1
  int64_t nnz;  // sparse only
};

constexpr int PW_MAX_OPS = 48;        // longest chain of the persistent form
constexpr int PW_PF = 4;              // entries per row loaded ahead
constexpr int PW_RP_LDS = 96 * 1024;  // dynamic LDS for cached row pointers (bytes)

__device__ __forceinline__ int pw_sgpr(int x) { return __builtin_amdgcn_readfirstlane(x); }

// canonical reduction of dot(x, y) over n values held in LDS; every thread
// returns the same value.  red: 256 doubles of LDS scratch; tot: 1 double.
// Per 4096-value chunk: lane t (< 256) sums x[i] y[i] for i = t, t + 256, ...
// in order, then the halving tree red[t] += red[t + h], h = 128 ... 1; the
// last six levels run inside wave 0 through lane shuffles (the same pairs,
// in the same order).  Chunk sums are added in chunk order.
__device__ double pw_dot(const double* x, const double* y, int n, double* red, double* tot) {
  const int t = threadIdx.x;
  const int nch = (n + RED_CHUNK - 1) / RED_CHUNK;
  double acc_tot = 0.0;
  for (int c = 0; c < nch; ++c) {
    if (t < RED_LANES) {
      double acc = 0.0;
      for (int m = 0; m < RED_CHUNK / RED_LANES; ++m) {
        int i = c * RED_CHUNK + m * RED_LANES + t;
        if (i < n) acc = fma(x[i], y[i], acc);
      }
      red[t] = acc;
    }
    __syncthreads();
    if (t < 128) red[t] = red[t] + red[t + 128];
    __syncthreads();
    if (t < 64) {
      double v = red[t] + red[t + 64];
#pragma unroll
      for (int h = 32; h >= 1; h >>= 1) v = v + __shfl_down(v, h, 64);
      if (t == 0) acc_tot = acc_tot + v;
    }
  }
  if (t == 0) *tot = acc_tot;
  __syncthreads();
  double r = *tot;
  __syncthreads();
  // the same on every lane; said so to the compiler, which then keeps the
  // loop tests that depend on it (and everything inside) scalar
  const unsigned long long u = __builtin_bit_cast(unsigned long long, r);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u), hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Entries of one row of a sparse operator, loaded ahead of their use.
struct PwPre {
  double v[PW_PF];
  int c[PW_PF];
  int p, n;  // first entry and count; n > PW_PF: read in the loop
};

typedef unsigned int pw_u4 __attribute__((ext_vector_type(4)));
typedef double pw_d2 __attribute__((ext_vector_type(2)));
constexpr int PW_RSRC_WORD3 = 0x00020000;

// Global-memory views of the operator arrays for the loops over long rows and
// dense operators (vector memory loads, counted on vmcnt only).
typedef const __attribute__((address_space(1))) double* pw_gd;
typedef const __attribute__((address_space(1))) int* pw_gi;
// Operator descriptors: read-only for the whole launch, read word by word
// through the constant address space (scalar loads).
typedef const __attribute__((address_space(4))) int* pw_cop;
struct PwOpWords {
  int w[sizeof(PwOp) / 4];
};
static_assert(sizeof(PwOp) % 4 == 0, "PwOp is read as words");
__device__ __forceinline__ PwOp pw_op(pw_cop ops, int s) {
  PwOpWords x;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(PwOp) / 4); ++i) x.w[i] = ops[s * (int)(sizeof(PwOp) / 4) + i];
  return __builtin_bit_cast(PwOp, x);
}

// Barrier between operators: orders the LDS vectors only, without waiting
// for the entry loads of the next operators.
__device__ __forceinline__ void pw_lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Loads the first PW_PF entries of row r: two 16-byte buffer loads of values
// and one of column indices, whatever the row and the operator.  The buffers
// cover exactly the operator's entries (none for a dense operator), so reads
// past them return zeros; rows past the operator and dense operators get
// n = 0.  Every thread thus issues the same three loads for every operator,
// and the wait before their use covers only the loads issued before them.
// RP_LDS: the row pointers are in LDS (a dense operator's rpo points at valid
// LDS), else they are read from global memory.
// The descriptor fields pw_fetch needs, read from LDS into scalar registers
// (the same for every lane), so the buffer resources are scalar.
struct PwU {
  const double* v;
  const int* ci;
  const int* rp;
  int kind, rows, rpo, nnz;
};
template <class T>
__device__ __forceinline__ T* pw_sptr(T* p) {
  const unsigned long long u = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u), hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return (T*)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ PwU pw_uniform(const PwOp o) {
  PwU u;
  u.v = pw_sptr(o.v);
  u.ci = pw_sptr(o.ci);
  u.rp = pw_sptr(o.rp);
  u.kind = pw_sgpr(o.kind);
  u.rows = pw_sgpr(o.rows);
  u.rpo = pw_sgpr(o.rpo);
  u.nnz = pw_sgpr((int)o.nnz);
  return u;
}

template <bool RP_LDS, bool SHORT = false>
__device__ __forceinline__ void pw_fetch(const PwOp od, const int* rpl, int r, PwPre& f) {
  const PwU o = pw_uniform(od);
  const bool sp = SHORT || o.kind == 1;
  const int rr = min(r, o.rows - 1);
  int p0, p1;
  if (RP_LDS) {
    p0 = rpl[o.rpo + rr];
    p1 = rpl[o.rpo + rr + 1];
  } else {
    const pw_gi rp = (pw_gi)o.rp;
    p0 = sp ? rp[rr] : 0;
    p1 = sp ? rp[rr + 1] : 0;
  }
  f.p = p0;
  f.n = (sp && r < o.rows) ? p1 - p0 : 0;
  const int vb = sp ? o.nnz * 8 : 0, cb = sp ? o.nnz * 4 : 0;
  const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc((void*)o.v, 0, vb, PW_RSRC_WORD3);
  const __amdgpu_buffer_rsrc_t rc = __builtin_amdgcn_make_buffer_rsrc((void*)o.ci, 0, cb, PW_RSRC_WORD3);
  const pw_d2 v01 = __builtin_bit_cast(pw_d2, __builtin_amdgcn_raw_buffer_load_b128(rv, p0 * 8, 0, 0));
  const pw_d2 v23 = __builtin_bit_cast(pw_d2, __builtin_amdgcn_raw_buffer_load_b128(rv, p0 * 8 + 16, 0, 0));
  const pw_u4 c = __builtin_amdgcn_raw_buffer_load_b128(rc, p0 * 4, 0, 0);
  f.v[0] = v01.x;
  f.v[1] = v01.y;
  f.v[2] = v23.x;
  f.v[3] = v23.y;
  f.c[0] = (int)c.x;
  f.c[1] = (int)c.y;
  f.c[2] = (int)c.z;
  f.c[3] = (int)c.w;
}

// row r of op(o) applied to cur (LDS), entries from f when they were loaded.
// SHORT: every operator of the chain is sparse with rows of at most PW_PF
// entries, so the row is the prefetched slots alone, taken without branches:
// every slot's column is a valid column of the operator (an entry of this
// row or a later one, or 0 past the entries) and the terms past the row end
// are computed and not kept (the same fma chain, the same bits).
// K (SHORT only): the operator's longest row, the same on every lane, so that
// rows of at most two entries gather two slots, not four.
template <bool SHORT>
__device__ __forceinline__ double pw_row(const PwOp o, const PwPre& f, int r, const double* cur, int K = PW_PF) {
  double acc = 0.0;
  if (SHORT) {
    // two slots when no row has more than two entries, else all four; the
    // gathers of a case are issued before its fma chain
    if (K <= 2) {
      const double x0 = cur[f.c[0]], x1 = cur[f.c[1]];
      const double t0 = fma(f.v[0], x0, acc);
      acc = (0 < f.n) ? t0 : acc;
      const double t1 = fma(f.v[1], x1, acc);
      return (1 < f.n) ? t1 : acc;
    }
    double x[PW_PF];
#pragma unroll
    for (int k = 0; k < PW_PF; ++k) x[k] = cur[f.c[k]];
#pragma unroll
    for (int k = 0; k < PW_PF; ++k) {
      const double tk = fma(f.v[k], x[k], acc);
      acc = (k < f.n) ? tk : acc;
    }
    return acc;
  }
  if (o.kind == 1) {
    if (f.n <= PW_PF) {
#pragma unroll
      for (int k = 0; k < PW_PF; ++k)
        if (k < f.n) acc = fma(f.v[k], cur[f.c[k]], acc);
    } else {
      const pw_gd v = (pw_gd)o.v;
      const pw_gi ci = (pw_gi)o.ci;
      for (int p = f.p, pe = f.p + f.n; p < pe; ++p) acc = fma(v[p], cur[ci[p]], acc);
    }
  } else if (!o.trans) {
    const pw_gd a = (pw_gd)o.v + r;
    for (int kk = 0; kk < o.cols; ++kk) acc = fma(a[(size_t)kk * o.lda], cur[kk], acc);
  } else {
    const pw_gd a = (pw_gd)o.v + (size_t)r * o.lda;
    for (int kk = 0; kk < o.cols; ++kk) acc = fma(a[kk], cur[kk], acc);
  }
  return acc;
}

// One operator of the chain: row t (and t + PW_THREADS when TWO_ROWS) of
// operator s from cur into nxt with the entries in (f0, f1), which are then
// reloaded with those of operator s - 2; ends with the LDS barrier.
// Marks the prefetched registers as read on every path, so that the wait for
// their loads sits here (with the next operator's loads still in flight) and
// the registers can be reused afterwards without a wait for everything.
__device__ __forceinline__ void pw_settle(const PwPre& f) {
  asm volatile("" ::"v"(f.v[0]), "v"(f.v[1]), "v"(f.v[2]), "v"(f.v[3]), "v"(f.c[0]), "v"(f.c[1]), "v"(f.c[2]),
               "v"(f.c[3]));
}

template <bool RP_LDS, bool TWO_ROWS, bool SHORT>
__device__ __forceinline__ void pw_apply(pw_cop sops, const int* rpl, const int* opw, int s, const double* cur,
                                         double* nxt, PwPre& f0, PwPre& f1) {
  const int t = threadIdx.x;
  pw_settle(f0);
  if (TWO_ROWS) pw_settle(f1);
  s = pw_sgpr(s);  // the same on every lane: descriptors come through scalar loads
  const PwOp o = pw_op(sops, s);
  const int rows = o.rows;
  const int K = SHORT ? pw_sgpr(opw[s]) : PW_PF;
  double y0 = 0.0, y1 = 0.0;
  if (t < rows) y0 = pw_row<SHORT>(o, f0, t, cur, K);
  if (TWO_ROWS && t + PW_THREADS < rows) y1 = pw_row<SHORT>(o, f1, t + PW_THREADS, cur, K);
  // (operators 1 and 0 reload operator 0: unused, keeps the loads unconditional)
  const PwOp o2 = pw_op(sops, max(s - 2, 0));
  pw_fetch<RP_LDS, SHORT>(o2, rpl, t, f0);
  if (TWO_ROWS) pw_fetch<RP_LDS, SHORT>(o2, rpl, t + PW_THREADS, f1);
  if (t < rows) nxt[t] = y0;
  if (TWO_ROWS && t + PW_THREADS < rows) nxt[t + PW_THREADS] = y1;
  pw_lds_barrier();
}

// out[0] = lambda, out[1] = flag (iterations, or -1), out[2] = status
// (0 ok, 1 zero-norm vector, 2 A x == 0)
// One chain per workgroup: workgroup b runs jobs[b] (independent chains, e.g.
// the L and R norms of one PALM factor update, share a launch).  A chain has
// at most PW_MAX_OPS operators.
struct PwJob {
  const PwOp* ops;
  int nops;
  int n;
  const double* x0;
  double* out;
};

// The rounds of one chain (the reference's loop, faust_linear_algebra.hpp:
// 808-898): normalise xa into xn, apply the operators last to first, Rayleigh
// quotient.  Instantiated for SHORT chains (every operator sparse, rows of at
// most PW_PF entries: no loops over entries, so nothing but the prefetch is
// in flight across the operator barriers) and for the general case.
template <bool RP_LDS, bool TWO_ROWS, bool SHORT>
__device__ __forceinline__ void pw_rounds(pw_cop sops, const int* rpl, const int* opw, int nops, int n, int maxit,
                                          double thr,
                                          double* xa, double* xb, double* xn, double* red, double* tot,
                                          double& lambda, int& it, int& status) {
  const int t = threadIdx.x;
  double lambda_old = 1.0;
  const double athr = fabs(thr);
  // entries of rows t and t + PW_THREADS (rows <= PW_VMAX = 2 x PW_THREADS) of
  // the operators applied 1st, 3rd, ... (a) and 2nd, 4th, ... (b)
  PwPre a0, a1, b0, b1;
  while ((fabs(lambda_old - lambda) > athr || fabs(lambda) <= athr) && it < maxit) {
    it++;
    lambda_old = lambda;
    double nrm = sqrt(pw_dot(xa, xa, n, red, tot));
    if (nrm == 0.0) {
      status = 1;
      break;
    }
    pw_fetch<RP_LDS, SHORT>(pw_op(sops, nops - 1), rpl, t, a0);
    if (TWO_ROWS) pw_fetch<RP_LDS, SHORT>(pw_op(sops, nops - 1), rpl, t + PW_THREADS, a1);
    pw_fetch<RP_LDS, SHORT>(pw_op(sops, max(nops - 2, 0)), rpl, t, b0);
    if (TWO_ROWS) pw_fetch<RP_LDS, SHORT>(pw_op(sops, max(nops - 2, 0)), rpl, t + PW_THREADS, b1);
    const double inv = 1 / nrm;
    for (int i = t; i < n; i += PW_THREADS) xn[i] = xa[i] * inv;
    pw_lds_barrier();
    // apply ops last to first, ping-pong between xb and xa (xn stays the
    // normalised start vector)
    const double* cur = xn;
    double* nxt = xb;
    int len = n;
    for (int s = nops - 1; s >= 0; s -= 2) {
      pw_apply<RP_LDS, TWO_ROWS, SHORT>(sops, rpl, opw, s, cur, nxt, a0, a1);
      len = pw_op(sops, s).rows;
      cur = nxt;
      nxt = (nxt == xb) ? xa : xb;
      if (s == 0) break;
      pw_apply<RP_LDS, TWO_ROWS, SHORT>(sops, rpl, opw, s - 1, cur, nxt, b0, b1);
      len = pw_op(sops, s - 1).rows;
      cur = nxt;
      nxt = (nxt == xb) ? xa : xb;
    }
    // result vector is `cur`; move it into xa if needed
    if (cur != xa) {
      for (int i = t; i < len; i += PW_THREADS) xa[i] = cur[i];
      __syncthreads();
    }
    int nz = 0;
    for (int i = t; i < n; i += PW_THREADS) nz |= (xa[i] != 0.0);
    if (!__syncthreads_or(nz)) {
      status = 2;
      break;
    }
    lambda = pw_dot(xn, xa, n, red, tot);
  }
}

template <bool RP_LDS, bool TWO_ROWS>
__global__ __launch_bounds__(PW_THREADS) void k_power_iteration(const PwJob* __restrict__ jobs, int maxit,
                                                                 double thr) {
  extern __shared__ int rpl[];  // cached row pointers (PwOp::rpo)
  const PwJob job = jobs[blockIdx.x];
  const int nops = job.nops, n = job.n;
  const double* __restrict__ x0 = job.x0;
  double* __restrict__ out = job.out;
  __shared__ double xa[PW_VMAX], xb[PW_VMAX], xn[PW_VMAX];
  __shared__ double red[RED_LANES];
  __shared__ double tot;
  // operator descriptors: read-only and indexed by loop counters, so they
  // come through scalar loads
  const pw_cop sops = (pw_cop)job.ops;
  const int t = threadIdx.x;
  for (int i = t; i < n; i += PW_THREADS) xa[i] = x0[i];
  // copy the row pointers, find each operator's longest row (opw) and whether
  // every operator is sparse with rows of at most PW_PF entries
  __shared__ int opw[PW_MAX_OPS];
  for (int s = t; s < nops; s += PW_THREADS) opw[s] = 0;
  __syncthreads();
  int wide_row = 0;
  for (int s = 0; s < nops; ++s) {
    const PwOp o = pw_op(sops, s);
    if (o.kind != 1) {
      wide_row = 1;
      continue;
    }
    int w = 0;
    for (int i = t; i <= o.rows; i += PW_THREADS) {
      const int q = o.rp[i];
      if (RP_LDS) rpl[o.rpo + i] = q;
      if (i < o.rows) w = max(w, o.rp[i + 1] - q);
    }
    wide_row |= w > PW_PF;
    if (w > 0) atomicMax(&opw[s], w);
  }
  const bool short_chain = !__syncthreads_or(wide_row);
  double lambda = 0.0;
  int it = 0;
  int status = 0;
  if (short_chain)
    pw_rounds<RP_LDS, TWO_ROWS, true>(sops, rpl, opw, nops, n, maxit, thr, xa, xb, xn, red, &tot, lambda, it,
                                      status);
  else
    pw_rounds<RP_LDS, TWO_ROWS, false>(sops, rpl, opw, nops, n, maxit, thr, xa, xb, xn, red, &tot, lambda, it,
                                       status);
  if (t == 0) {
    out[0] = (status == 2) ? 0.0 : lambda;
    out[1] = (it < maxit) ? (double)it : -1.0;
    out[2] = (double)status;
  }
}

// ---------------------------------------------------------------------------
// Wide form: one launch per operator per round, for chains whose operators are
// too large for a single CU to stream every round.  Each row is still one fma
// chain on one lane (k ascending), so the result is bit-identical to the
// persistent kernel; the row data is staged through LDS by all 256 threads of
// the workgroup in contiguous chunks, and the first R lanes (one per row) run
// the chains.  A device state carries lambda, the round count and a done flag: the
// host enqueues rounds ahead and launches after convergence return at once.
struct PwState {
  double lambda, lambda_old;
  int it, status, done, pad;
};

constexpr int PW_WIDE_THREADS = 256;
constexpr int PW_CH = 2048;  // entries per LDS chunk

// sparse rows (row pointers rp, indices ci, values v): rows r0..r0+R-1 per workgroup
__global__ __launch_bounds__(PW_WIDE_THREADS) void k_pw_csr(const int* __restrict__ rp, const int* __restrict__ ci,
                                                            const double* __restrict__ v, int rows, int cols, int R,
                                                            const double* __restrict__ cur, double* __restrict__ nxt,
                                                            const PwState* __restrict__ s) {
  __shared__ double xs[PW_VMAX];
  __shared__ double sv[PW_CH];
  __shared__ int sci[PW_CH];
  if (s->done) return;
  const int t = threadIdx.x;
  const int r0 = blockIdx.x * R, r1 = min(rows, r0 + R), nr = r1 - r0;
  for (int i = t; i < cols; i += PW_WIDE_THREADS) xs[i] = cur[i];
  const int p0 = rp[r0], p1 = rp[r1];
  int p = 0, pe = 0;
  if (t < nr) {
    p = rp[r0 + t];
    pe = rp[r0 + t + 1];
  }
  double acc = 0.0;
  for (int c0 = p0; c0 < p1; c0 += PW_CH) {
    const int c1 = min(p1, c0 + PW_CH);
    __syncthreads();  // previous chunk consumed; xs staged before the first
    for (int q = c0 + t; q < c1; q += PW_WIDE_THREADS) {
      sv[q - c0] = v[q];
      sci[q - c0] = ci[q];
    }
    __syncthreads();
    if (t < nr) {
      const int e = min(pe, c1);
      for (; p < e; ++p) acc = fma(sv[p - c0], xs[sci[p - c0]], acc);
    }
  }
  if (t < nr) nxt[r0 + t] = acc;
}

// dense op(A) rows (A column-major, leading dimension lda); trans: op(A) = A^T
__global__ __launch_bounds__(PW_WIDE_THREADS) void k_pw_dense(const double* __restrict__ A, int lda, int trans,
                                                              int rows, int cols, int R,
                                                              const double* __restrict__ cur,
                                                              double* __restrict__ nxt,
                                                              const PwState* __restrict__ s) {
  __shared__ double xs[PW_VMAX];
  __shared__ double tile[PW_CH];  // R rows x C columns, row stride C + 1 (odd: no bank conflicts)
  if (s->done) return;
  const int t = threadIdx.x;
  const int r0 = blockIdx.x * R, nr = min(rows, r0 + R) - r0;
  const int C = PW_CH / R - 2;
  for (int i = t; i < cols; i += PW_WIDE_THREADS) xs[i] = cur[i];
  double acc = 0.0;
  for (int c0 = 0; c0 < cols; c0 += C) {
    const int nc = min(C, cols - c0);
    __syncthreads();
    if (trans) {  // row i of A^T = column r0+i of A: contiguous in c
      for (int q = t; q < nr * nc; q += PW_WIDE_THREADS) {
        const int i = q / nc, c = q - i * nc;
        tile[i * (C + 1) + c] = A[(size_t)(r0 + i) * lda + c0 + c];
      }
    } else {  // row i of A: element (i, c) at (c0+c)*lda + r0 + i
      for (int q = t; q < nr * nc; q += PW_WIDE_THREADS) {
        const int c = q / nr, i = q - c * nr;
        tile[i * (C + 1) + c] = A[(size_t)(c0 + c) * lda + r0 + i];
      }
    }
    __syncthreads();
    if (t < nr) {
      const double* row = tile + t * (C + 1);
      for (int c = 0; c < nc; ++c) acc = fma(row[c], xs[c0 + c], acc);
    }
  }
  if (t < nr) nxt[r0 + t] = acc;
}

// Operators with more than PW_VMAX columns: x read from global memory (the
// L2 holds it), one row per thread, the same fma chain per row.
__global__ __launch_bounds__(PW_WIDE_THREADS) void k_pw_csr_g(const int* __restrict__ rp, const int* __restrict__ ci,
                                                              const double* __restrict__ v, int rows,
                                                              const double* __restrict__ cur,
                                                              double* __restrict__ nxt,
                                                              const PwState* __restrict__ s) {
  if (s->done) return;
  const int r = blockIdx.x * PW_WIDE_THREADS + threadIdx.x;
  if (r >= rows) return;
  double acc = 0.0;
  for (int p = rp[r], pe = rp[r + 1]; p < pe; ++p) acc = fma(v[p], cur[ci[p]], acc);
  nxt[r] = acc;
}

__global__ __launch_bounds__(PW_WIDE_THREADS) void k_pw_dense_g(const double* __restrict__ A, int lda, int trans,
                                                                int rows, int cols, const double* __restrict__ cur,
                                                                double* __restrict__ nxt,
                                                                const PwState* __restrict__ s) {
  if (s->done) return;
  const int r = blockIdx.x * PW_WIDE_THREADS + threadIdx.x;
  if (r >= rows) return;
  double acc = 0.0;
  if (trans) {  // row r of A^T = column r of A
    const double* col = A + (size_t)r * lda;
    for (int c = 0; c < cols; ++c) acc = fma(col[c], cur[c], acc);
  } else {
    for (int c = 0; c < cols; ++c) acc = fma(A[(size_t)c * lda + r], cur[c], acc);
  }
  nxt[r] = acc;
}

// One round boundary.  mode 1: finish the previous round (xa = chain output):
// zero test and lambda = <xn, xa>.  Then the loop test of the persistent
// kernel, and if it continues: it++, normalise xa into xn.
__global__ __launch_bounds__(PW_THREADS) void k_pw_step(int n, double* __restrict__ xa, double* __restrict__ xn,
                                                         PwState* __restrict__ s, int maxit, double thr, int mode) {
  __shared__ double red[RED_LANES];
  __shared__ double tot;
  const int t = threadIdx.x;
  const PwState st = *s;
  __syncthreads();
  if (st.done) return;
  double lambda = st.lambda, lambda_old = st.lambda_old;
  int it = st.it;
  if (mode == 1) {
    int nz = 0;
    for (int i = t; i < n; i += PW_THREADS) nz |= (xa[i] != 0.0);
    if (!__syncthreads_or(nz)) {
      if (t == 0) {
        s->status = 2;
        s->done = 1;
      }
      return;
    }
    lambda = pw_dot(xn, xa, n, red, &tot);
  }
  const double athr = fabs(thr);
  if (!((fabs(lambda_old - lambda) > athr || fabs(lambda) <= athr) && it < maxit)) {
    if (t == 0) {
      s->lambda = lambda;
      s->done = 1;
    }
    return;
  }
  it++;
  lambda_old = lambda;
  const double nrm = sqrt(pw_dot(xa, xa, n, red, &tot));
  if (nrm == 0.0) {
    if (t == 0) {
      s->it = it;
      s->status = 1;
      s->done = 1;
    }
    return;
  }
  const double inv = 1 / nrm;
  for (int i = t; i < n; i += PW_THREADS) xn[i] = xa[i] * inv;
  if (t == 0) {
    s->lambda = lambda;
    s->lambda_old = lambda_old;
    s->it = it;
  }
}

__global__ void k_pw_init(PwState* __restrict__ s) {
  if (threadIdx.x == 0) {
    s->lambda = 0.0;
    s->lambda_old = 1.0;
    s->it = 0;
    s->status = 0;
    s->done = 0;
    s->pad = 0;
  }
}

const std::vector<double>& power_x0(int n) {
  static std::mutex mu;
  static std::map<int, std::vector<double>> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(n);
  if (it != cache.end()) return it->second;
  // srand(0xF4+'u'+57); Eigen::Matrix::Random: x + (y-x)*Scalar(std::rand())/Scalar(RAND_MAX) on [-1, 1]
  srand(0xF4 + 'u' + 57);
  std::vector<double> x(n);
  for (int i = 0; i < n; ++i) x[i] = -1.0 + (2.0 * (double)std::rand()) / (double)RAND_MAX;
  return cache.emplace(n, std::move(x)).first->second;
}

// Per-round time estimates (seconds): the persistent kernel streams every
// operator through one CU (~100 GB/s); the wide form pays a launch per
// operator but spreads the rows over the chip, leaving each row's fma chain
// (about 4 ns per term) as its floor.  FH_POWER_WIDE=0/1 forces a form.
static bool use_wide_form(const std::vector<PwOp>& h) {
  if (const char* e = getenv("FH_POWER_WIDE")) {
    if (e[0] == '0') return false;
    if (e[0] == '1') return true;
  }
  double single = 3e-6, wide = 4e-6;
  for (const PwOp& o : h) {
    const double entries = o.kind == 1 ? (double)o.nnz : (double)o.rows * o.cols;
    const double bytes = entries * (o.kind == 1 ? 12.0 : 8.0);
    single += bytes / 100e9 + 0.4e-6;
    wide += 3e-6 + (entries / std::max(o.rows, 1)) * 4e-9 + bytes / 1.5e12;
  }
  return wide < single;
}

static double power_iteration_wide(const std::vector<PwOp>& h, int n, const std::vector<double>& x0, int maxit,
                                   double thr, int* flag, hipStream_t st) {
  int vmax = n;
  for (const PwOp& o : h) vmax = std::max(vmax, std::max(o.rows, o.cols));
  DBuf xa(sizeof(double) * n), xn(sizeof(double) * n), b0(sizeof(double) * vmax), b1(sizeof(double) * vmax),
      ds(sizeof(PwState));
  PwState* dst = ds.as<PwState>();
  FH_HIP(hipMemcpyAsync(xa.p, x0.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_pw_init, dim3(1), dim3(64), 0, st, dst);
  FH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_pw_step, dim3(1), dim3(PW_THREADS), 0, st, n, xa.as<double>(), xn.as<double>(), dst, maxit,
                     thr, 0);
  FH_LAUNCH_CHECK();
  const int nops = (int)h.size();
  PwState hs{};
  int next_check = 4;
  for (int round = 1; round <= maxit; ++round) {
    const double* cur = xn.as<double>();
    for (int s = nops - 1; s >= 0; --s) {
      const PwOp& o = h[s];
      double* out = (s == 0) ? xa.as<double>() : (((nops - 1 - s) & 1) ? b1.as<double>() : b0.as<double>());
      int R = 1;
      while (R < 8 && (o.rows + 2 * R - 1) / (2 * R) >= 256) R *= 2;
      const dim3 grid((o.rows + R - 1) / R);
      if (o.cols > PW_VMAX) {  // x does not fit LDS: read it from global memory
        const dim3 gg((o.rows + PW_WIDE_THREADS - 1) / PW_WIDE_THREADS);
        if (o.kind == 1)
          hipLaunchKernelGGL(k_pw_csr_g, gg, dim3(PW_WIDE_THREADS), 0, st, o.rp, o.ci, o.v, o.rows, cur, out,
                             (const PwState*)dst);
        else
          hipLaunchKernelGGL(k_pw_dense_g, gg, dim3(PW_WIDE_THREADS), 0, st, o.v, o.lda, o.trans, o.rows, o.cols,
                             cur, out, (const PwState*)dst);
      } else if (o.kind == 1)
        hipLaunchKernelGGL(k_pw_csr, grid, dim3(PW_WIDE_THREADS), 0, st, o.rp, o.ci, o.v, o.rows, o.cols, R, cur,
                           out, (const PwState*)dst);
      else
        hipLaunchKernelGGL(k_pw_dense, grid, dim3(PW_WIDE_THREADS), 0, st, o.v, o.lda, o.trans, o.rows, o.cols, R,
                           cur, out, (const PwState*)dst);
      FH_LAUNCH_CHECK();
      cur = out;
    }
    hipLaunchKernelGGL(k_pw_step, dim3(1), dim3(PW_THREADS), 0, st, n, xa.as<double>(), xn.as<double>(), dst,
                       maxit, thr, 1);
    FH_LAUNCH_CHECK();
    if (round == next_check || round == maxit) {
      FH_HIP(hipMemcpyAsync(&hs, dst, sizeof(PwState), hipMemcpyDeviceToHost, st));
      FH_HIP(hipStreamSynchronize(st));
      if (hs.done) break;
      next_check += 8;
    }
  }
  if (!hs.done) {  // every round ran: one more step evaluates the loop test
    hipLaunchKernelGGL(k_pw_step, dim3(1), dim3(PW_THREADS), 0, st, n, xa.as<double>(), xn.as<double>(), dst,
                       maxit, thr, 1);
    FH_LAUNCH_CHECK();
    FH_HIP(hipMemcpyAsync(&hs, dst, sizeof(PwState), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
  }
  if (hs.status == 1) throw Error(FH_ERR_NUMERIC, "Can't normalize a zero-norm vector.");
  if (flag) *flag = (hs.it < maxit) ? hs.it : -1;
  return hs.status == 2 ? 0.0 : hs.lambda;
}

static std::vector<PwOp> pw_ops(const std::vector<ChainOp>& ops) {
  FH_CHECK(!ops.empty(), "power_iteration: empty chain");
  std::vector<PwOp> h(ops.size());
  for (size_t s = 0; s < ops.size(); ++s) {
    const DMat& M = *ops[s].M;
    const bool tr = ops[s].op != 'N';
    PwOp& o = h[s];
    o.rpo = -1;
    o.rows = tr ? M.n : M.m;
    o.cols = tr ? M.m : M.n;
    if (M.kind == K_CSR) {
      FH_CHECK(M.dtype == FH_F64, "power_iteration: FP64 dense or CSR factors only");
      o.kind = 1;
      o.trans = 0;
      o.nnz = M.nnz;
      if (tr) {
        FH_CHECK(M.has_csc, "power_iteration: CSC twin missing");
        o.rp = M.cp.as<int>();
        o.ci = M.ri.as<int>();
        o.v = M.cv.as<double>();
      } else {
        o.rp = M.rp.as<int>();
        o.ci = M.ci.as<int>();
        o.v = M.v.as<double>();
      }
    } else {
      FH_CHECK(M.kind == K_DENSE && M.dtype == FH_F64, "power_iteration: FP64 dense or CSR factors only");
      o.kind = 0;
      o.trans = tr ? 1 : 0;
      o.v = M.d.as<double>();
      // read (and unused) by the persistent kernel's unconditional loads
      o.rp = o.ci = reinterpret_cast<const int*>(o.v);
      o.lda = M.m;
      o.nnz = 0;
    }
  }
  return h;
}

// the persistent kernel keeps every vector in LDS: up to PW_VMAX entries
static bool pw_persistent(const std::vector<PwOp>& h, int n) {
  bool fits = n <= PW_VMAX && h.size() <= (size_t)PW_MAX_OPS;
  for (const PwOp& o : h) fits = fits && o.rows >= 1 && o.rows <= PW_VMAX && o.cols <= PW_VMAX;
  return fits && !use_wide_form(h);
}

// Dynamic LDS the persistent kernel may take for row pointers: PW_RP_LDS
// once the attribute is granted, else none (every operator then reads its
// row pointers from global memory).  Queried once per process.
static size_t pw_rp_budget() {
  static int budget = -1;
  if (budget < 0)
    budget = hipFuncSetAttribute((const void*)k_power_iteration<true, false>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, PW_RP_LDS) == hipSuccess &&
                     hipFuncSetAttribute((const void*)k_power_iteration<true, true>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, PW_RP_LDS) == hipSuccess
                 ? PW_RP_LDS
                 : 0;
  (void)hipGetLastError();
  return (size_t)budget;
}

// Row-pointer copies of a chain's sparse operators in LDS (an array shared by
// two operators is copied once); returns the ints used, or -1 (and no
// copies) when they do not all fit the budget.
static long pw_assign_rp(std::vector<PwOp>& ops, size_t budget_ints) {
  size_t used = 0, maxrows = 0;
  for (size_t s = 0; s < ops.size(); ++s) {
    PwOp& o = ops[s];
    o.rpo = -1;
    maxrows = std::max(maxrows, (size_t)o.rows);
    if (o.kind != 1) {
      o.rpo = 0;  // a dense operator's (unused) row-pointer reads stay inside the array
      continue;
    }
    for (size_t q = 0; q < s; ++q)
      if (ops[q].kind == 1 && ops[q].rp == o.rp) o.rpo = ops[q].rpo;
    if (o.rpo >= 0) continue;
    o.rpo = (int)used;
    used += o.rows + 1;
  }
  used = std::max(used, maxrows + 1);
  if (used > budget_ints) {
    for (PwOp& o : ops) o.rpo = -1;
    return -1;
  }
  return (long)used;
}

// chains[k] (ops, n) in one launch, one workgroup each; lam[k] and flags[k]
// receive each chain's result.  A zero-norm start vector throws, for the
// first such chain in order.
static void power_iteration_persistent(const std::vector<std::vector<PwOp>>& hs_in, const std::vector<int>& ns,
                                       int maxit, double thr, double* lam, int* flags, hipStream_t st) {
  const int J = (int)hs_in.size();
  std::vector<std::vector<PwOp>> hs = hs_in;
  size_t nops = 0, nx = 0;
  long dyn_ints = 0;
  const size_t budget = pw_rp_budget() / sizeof(int);
  bool rp_lds = true;
  for (int k = 0; k < J; ++k) {
    nops += hs[k].size();
    nx += ns[k];
    const long used = pw_assign_rp(hs[k], budget);
    rp_lds = rp_lds && used >= 0;
    dyn_ints = std::max(dyn_ints, used);
  }
  if (!rp_lds) {
    for (auto& h : hs)
      for (PwOp& o : h) o.rpo = -1;
    dyn_ints = 0;
  }
  DBuf dops(sizeof(PwOp) * nops), dx(sizeof(double) * nx), dout(sizeof(double) * 3 * J), djobs(sizeof(PwJob) * J);
  std::vector<PwOp> allops;
  std::vector<PwJob> jobs(J);
  size_t oo = 0, ox = 0;
  for (int k = 0; k < J; ++k) {
    allops.insert(allops.end(), hs[k].begin(), hs[k].end());
    FH_HIP(hipMemcpyAsync(dx.as<double>() + ox, power_x0(ns[k]).data(), sizeof(double) * ns[k],
                          hipMemcpyHostToDevice, st));
    jobs[k] = PwJob{dops.as<PwOp>() + oo, (int)hs[k].size(), ns[k], dx.as<double>() + ox, dout.as<double>() + 3 * k};
    oo += hs[k].size();
    ox += ns[k];
  }
  FH_HIP(hipMemcpyAsync(dops.p, allops.data(), sizeof(PwOp) * nops, hipMemcpyHostToDevice, st));
  FH_HIP(hipMemcpyAsync(djobs.p, jobs.data(), sizeof(PwJob) * J, hipMemcpyHostToDevice, st));
  bool two = false;  // some vector or operator longer than one row per thread
  for (int k = 0; k < J; ++k) {
    two = two || ns[k] > PW_THREADS;
    for (const PwOp& o : hs[k]) two = two || o.rows > PW_THREADS;
  }
  const size_t dyn = rp_lds ? sizeof(int) * dyn_ints : 0;
  auto kern = rp_lds ? (two ? k_power_iteration<true, true> : k_power_iteration<true, false>)
                     : (two ? k_power_iteration<false, true> : k_power_iteration<false, false>);
  hipLaunchKernelGGL(kern, dim3(J), dim3(PW_THREADS), dyn, st, djobs.as<PwJob>(), maxit, thr);
  FH_LAUNCH_CHECK();
  std::vector<double> res(3 * J);
  FH_HIP(hipMemcpyAsync(res.data(), dout.p, sizeof(double) * 3 * J, hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < J; ++k) {
    if (res[3 * k + 2] == 1.0) throw Error(FH_ERR_NUMERIC, "Can't normalize a zero-norm vector.");
    lam[k] = res[3 * k];
    flags[k] = (int)res[3 * k + 1];
  }
}

double power_iteration_f64(const std::vector<ChainOp>& ops, int n, int maxit, double thr, int* flag,
                           hipStream_t st) {
  FH_CHECK(maxit > 0, "power_iteration: nbr_iter_max <= 0");
  std::vector<PwOp> h = pw_ops(ops);
  if (!pw_persistent(h, n)) return power_iteration_wide(h, n, power_x0(n), maxit, thr, flag, st);
  double lam;
  int fl;
  power_iteration_persistent({h}, {n}, maxit, thr, &lam, &fl, st);
  if (flag) *flag = fl;
  return lam;
}

void power_iteration_pair_f64(const std::vector<ChainOp>& ops1, int n1, const std::vector<ChainOp>& ops2, int n2,
                              int maxit, double thr, double* lam1, double* lam2, int* flag, hipStream_t st) {
  FH_CHECK(maxit > 0, "power_iteration: nbr_iter_max <= 0");
  std::vector<PwOp> h1 = pw_ops(ops1), h2 = pw_ops(ops2);
  if (!pw_persistent(h1, n1) || !pw_persistent(h2, n2)) {
    *lam1 = power_iteration_f64(ops1, n1, maxit, thr, flag, st);
    *lam2 = power_iteration_f64(ops2, n2, maxit, thr, flag, st);
    return;
  }
  double lam[2];
  int fl[2];
  power_iteration_persistent({h1, h2}, {n1, n2}, maxit, thr, lam, fl, st);
  *lam1 = lam[0];
  *lam2 = lam[1];
  if (flag) *flag = fl[1];
}

}  // namespace fh
