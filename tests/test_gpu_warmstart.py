"""Learned warm start on the GPU: what the trained network gains over the
reference initialisation (small butterflies, n = 16, so the test runs in
seconds).  The network is trained by the library's own MLP kernels on
factorizations made by FP64 PALM4MSA; PALM4MSA then runs from its
prediction and from the reference initialisation with the same iteration
budget."""
import numpy as np
import pytest
import torch

from faust_amd import warmstart as ws

pytestmark = pytest.mark.gpu


def test_warm_start_beats_reference_init(gpu):
    n, L = 16, 4
    rng = np.random.default_rng(0x5EED)
    masks = ws.butterfly_supports(n)
    train = [ws.butterfly_sample(n, rng, 1e-6) for _ in range(2048)]
    Y, _, lerr = ws.palm_labels([a for a, _ in train], [g for _, g in train], masks, num_its=5)
    assert np.median(lerr) < 1e-5  # the labels are factorizations of their targets
    X = torch.from_numpy(np.stack([ws.mlp_input(a) for a, _ in train])).cuda()
    mlp = ws.MLP(n * n, 128, 128, 2 * n * L, seed=1)
    losses = ws.train_mlp(mlp, X, torch.from_numpy(Y).cuda(), epochs=30, batch=256, lr=1e-3)
    assert losses[-1] < 0.8 * losses[0]  # the network learns (standardised labels: start near 1)
    hold = [ws.butterfly_sample(n, rng, 1e-6)[0] for _ in range(64)]
    inits, lam0 = ws.predicted_init(mlp, hold, masks)
    res = ws.compare_warm_start(hold, masks, inits, lam0, max_its=500, check_every=25, target_err=1e-4,
                                dtype="f64")
    # before any fallback, the predicted start ends closer to the targets than
    # the reference start with the same iteration budget, for most targets
    assert res["warm_median_err_before_fallback"] < 0.7 * res["ref_median_err"], res
    assert res["fallbacks"] < 32, res
    assert res["warm_median_err"] <= res["ref_median_err"]


def test_butterfly_factorization_as_initialiser(gpu):
    """The butterfly factorization of a (noisy) butterfly target is already a
    factorization of it on the warm-start masks: the product is within the
    noise of the target before any PALM iteration, and PALM4MSA started
    there reaches 1e-4 at its first check for every target."""
    n = 16
    rng = np.random.default_rng(0xB0F)
    masks = ws.butterfly_supports(n)
    hold = [ws.butterfly_sample(n, rng, 1e-6)[0] for _ in range(32)]
    inits, lam = ws.butterfly_init(hold, masks)
    for A, fs, l in zip(hold, inits, lam):
        for f, M in zip(fs, masks):
            assert not np.any(f[M == 0])
        P = np.eye(n)
        for f in fs:
            P = P @ f
        assert np.linalg.norm(l * P - A) <= 1e-5 * np.linalg.norm(A)
    res = ws.compare_warm_start(hold, masks, inits, lam, max_its=100, check_every=25, target_err=1e-4,
                                dtype="f64")
    assert res["warm_reached_before_fallback"] == len(hold), res
    assert res["warm_median_its"] <= 25, res
