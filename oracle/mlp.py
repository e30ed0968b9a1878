"""FP64 CPU restatement of the warm-start MLP (TEST INFRASTRUCTURE ONLY).

Mirrors the GPU network's arithmetic points: inputs, weights, activations
and backward signals are rounded to BF16 (round to nearest even) exactly
where the GPU feeds them to its BF16 matrix products; everything else is
FP64 here (FP32 on the GPU).  GELU is the exact erf form."""
import math

import numpy as np
from scipy.special import erf


def bf16(a):
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)


def forward_backward(params, X, T):
    """params: [(W, b)] x 3 (W in x out).  Returns (loss, Y, grads)."""
    Wb = [bf16(W) for W, _ in params]
    Xb = bf16(X)
    Z, H, Hb = [], [], []
    inp = Xb
    for l in range(3):
        z = inp @ Wb[l] + params[l][1]
        h = gelu(z) if l < 2 else z
        Z.append(z)
        H.append(h)
        Hb.append(bf16(h))
        inp = Hb[-1]
    Y = H[2]
    n = Y.size
    loss = np.sum((Y - T) ** 2) / n
    dZ = 2.0 * (Y - T) / n
    grads = [None] * 3
    for l in (2, 1, 0):
        if l < 2:
            dZ = dZ * gelu_grad(Z[l])
        dZb = bf16(dZ)
        inp = Xb if l == 0 else Hb[l - 1]
        grads[l] = (inp.T @ dZb, dZ.sum(axis=0))
        if l > 0:
            dZ = dZb @ Wb[l].T
    return loss, Y, grads
