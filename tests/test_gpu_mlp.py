"""Warm-start MLP kernels: forward values and gradients against the FP64 CPU
restatement (oracle/mlp.py) on the same BF16-rounded operands.  Tolerance:
relative Frobenius error 2e-2 (FP32 accumulation order and BF16 re-rounding
of intermediates are the only differences), typical values are ~1e-4."""
import numpy as np
import pytest
import torch

from oracle import mlp as ref
from faust_amd.warmstart import MLP

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("dims,B", [((256, 64, 64, 128), 16), ((65536, 512, 512, 4096), 4)])
def test_mlp_gradients_match_oracle(gpu, dims, B):
    rng = np.random.default_rng(0)
    net = MLP(*dims, seed=3)
    params = []
    for l in range(3):
        W, b = net.get(l)
        b = (0.1 * rng.standard_normal(b.shape)).astype(np.float32)
        net.set(l, W, b)
        params.append((W.astype(np.float64), b.astype(np.float64)))
    X = rng.standard_normal((B, dims[0])).astype(np.float32)
    T = rng.standard_normal((B, dims[3])).astype(np.float32)
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    Y = net.forward(Xd).cpu().numpy()
    loss, grads = net.gradients(Xd, Td)
    rloss, rY, rgrads = ref.forward_backward(params, X.astype(np.float64), T.astype(np.float64))
    assert _rel(Y, rY) < 1e-2
    assert abs(loss - rloss) < 1e-2 * rloss
    for l in range(3):
        assert _rel(grads[l][0], rgrads[l][0]) < 2e-2, l
        assert _rel(grads[l][1], rgrads[l][1]) < 2e-2, l


def test_mlp_training_reduces_loss(gpu):
    rng = np.random.default_rng(1)
    dims = (256, 64, 64, 32)
    net = MLP(*dims, seed=5)
    Wt = rng.standard_normal((256, 32)) / 16
    X = rng.standard_normal((512, 256)).astype(np.float32)
    T = (np.tanh(X @ Wt)).astype(np.float32)
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    first = net.train_step(Xd, Td, 1e-3)
    for _ in range(200):
        last = net.train_step(Xd, Td, 1e-3)
    assert last < 0.2 * first
