"""Learned warm start for butterfly PALM4MSA.

Targets are products of L = log2(n) butterfly factors; factor j couples
indices i and i ^ 2^j through a 2x2 block s * [[cos t, -sin t], [sin t, cos t]]
(t ~ U(0, 2 pi), s ~ U(0.5, 2)), plus i.i.d. Gaussian noise.  PALM4MSA
factorizes them with `supp` projectors on the butterfly supports.

An MLP (n^2 -> 512 -> 512 -> 2 n L, GELU; BF16 weights on MFMA, FP32
accumulation; libfaust_hip's own forward/backward/Adam kernels) is trained
on factorizations produced by the plain FP64 PALM4MSA and predicts the 2 n L
nonzero values of the factors.  PALM4MSA starts from the prediction and a
run falls back to the reference initialisation when its final error is
worse than the plain run's.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


def butterfly_supports(n):
    """0/1 masks of the L butterfly factors (row i: columns i and i ^ 2^j)."""
    L = int(np.log2(n))
    masks = []
    i = np.arange(n)
    for j in range(L):
        m = np.zeros((n, n))
        m[i, i] = 1.0
        m[i, i ^ (1 << j)] = 1.0
        masks.append(np.asfortranarray(m))
    return masks


def random_butterfly_factor(n, j, rng):
    B = np.zeros((n, n))
    for i in range(n):
        if i & (1 << j):
            continue
        i2 = i | (1 << j)
        t = rng.uniform(0.0, 2 * np.pi)
        s = rng.uniform(0.5, 2.0)
        c, sn = s * np.cos(t), s * np.sin(t)
        B[i, i], B[i, i2], B[i2, i], B[i2, i2] = c, -sn, sn, c
    return B


def butterfly_target(n, rng, noise=1e-6):
    """A = B_0 B_1 ... B_{L-1} + N(0, noise^2) entries."""
    L = int(np.log2(n))
    P = np.eye(n)
    for j in range(L):
        P = P @ random_butterfly_factor(n, j, rng)
    return np.asfortranarray(P + noise * rng.standard_normal((n, n)))


def pack_values(factors, masks):
    """The 2 n per-factor nonzero values, factor-major, row-major inside a factor."""
    out = []
    for F, M in zip(factors, masks):
        r, c = np.nonzero(M)  # row-major order (numpy nonzero scans rows)
        out.append(F[r, c])
    return np.concatenate(out)


def unpack_values(v, masks):
    out = []
    off = 0
    for M in masks:
        r, c = np.nonzero(M)
        F = np.zeros(M.shape)
        F[r, c] = v[off:off + len(r)]
        off += len(r)
        out.append(np.asfortranarray(F))
    return out


def identity_init(masks):
    return [np.asfortranarray(np.eye(M.shape[0]) * (np.diag(M) != 0)) for M in masks]


class MLP:
    """The warm-start network on the GPU (fh_mlp_*)."""

    def __init__(self, n_in, n_h1=512, n_h2=512, n_out=4096, seed=0):
        self.dims = (n_in, n_h1, n_h2, n_out)
        self.h = C.c_void_p()
        check(lib().fh_mlp_create(n_in, n_h1, n_h2, n_out, seed, C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib._lib is not None:
            _lib._lib.fh_mlp_free(self.h)

    def _stream(self):
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def forward(self, Xd):
        """Xd: (B, n_in) float32 CUDA tensor -> (B, n_out) float32 CUDA tensor."""
        import torch
        Y = torch.empty((Xd.shape[0], self.dims[3]), dtype=torch.float32, device="cuda")
        check(lib().fh_mlp_forward(self.h, C.c_void_p(Xd.data_ptr()), Xd.shape[0], C.c_void_p(Y.data_ptr()),
                                   self._stream()))
        return Y

    def train_step(self, Xd, Td, lr):
        loss = C.c_float()
        check(lib().fh_mlp_train_step(self.h, C.c_void_p(Xd.data_ptr()), C.c_void_p(Td.data_ptr()), Xd.shape[0],
                                      float(lr), C.byref(loss), self._stream()))
        return loss.value

    def gradients(self, Xd, Td):
        loss = C.c_float()
        check(lib().fh_mlp_gradients(self.h, C.c_void_p(Xd.data_ptr()), C.c_void_p(Td.data_ptr()), Xd.shape[0],
                                     C.byref(loss), self._stream()))
        return loss.value, [self.get(l, what=1) for l in range(3)]

    def get(self, layer, what=0):
        K, N = self.dims[layer], self.dims[layer + 1]
        W = np.empty((K, N), dtype=np.float32)
        b = np.empty(N, dtype=np.float32)
        check(lib().fh_mlp_get(self.h, layer, what, W.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p)))
        return W, b

    def set(self, layer, W, b):
        W = np.ascontiguousarray(W, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        check(lib().fh_mlp_set(self.h, layer, W.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p)))
