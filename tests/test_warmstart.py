"""Butterfly targets and the canonical form of their factorizations (CPU)."""
import numpy as np
import pytest

from faust_amd import warmstart as ws


def _product(lam, facs):
    P = np.eye(facs[0].shape[0])
    for f in facs:
        P = P @ f
    return lam * P


def test_butterfly_target_structure():
    n = 32
    rng = np.random.default_rng(0)
    masks = ws.butterfly_supports(n)
    assert len(masks) == 5
    for j, M in enumerate(masks):
        assert np.all(M.sum(axis=0) == 2) and np.all(M.sum(axis=1) == 2)
        assert M[0, 1 << j] == 1 and M[1 << j, 0] == 1
    facs = [ws.random_butterfly_factor(n, j, rng) for j in range(5)]
    for f, M in zip(facs, masks):
        assert np.all(f[M == 0] == 0)
        # 2x2 blocks s * rotation: columns of a block are orthogonal with equal norms
        g = f.T @ f
        assert np.allclose(g, np.diag(np.diag(g)))
    v = ws.pack_values(facs, masks)
    assert v.shape == (2 * n * 5,)
    back = ws.unpack_values(v, masks)
    for a, b in zip(back, facs):
        assert np.array_equal(a, b)


def test_canonical_form_is_invariant_to_diagonal_rescaling():
    # F_j D_j, D_j^{-1} F_{j+1} with arbitrary diagonal D (signs included)
    # gives the same product and must give the same canonical factors
    n = 16
    rng = np.random.default_rng(1)
    facs = [ws.random_butterfly_factor(n, j, rng) for j in range(4)]
    lam0, can0 = ws.canonicalize(facs)
    assert np.allclose(_product(lam0, can0), _product(1.0, facs), rtol=0, atol=1e-12)
    G = [f.copy() for f in facs]
    for j in range(3):
        d = rng.uniform(0.3, 3.0, n) * rng.choice([-1.0, 1.0], n)
        G[j] = G[j] * d[None, :]
        G[j + 1] = G[j + 1] / d[:, None]
    G[0] *= 2.5
    G[2] *= 0.4  # 2.5 * 0.4 = 1: same product
    lam1, can1 = ws.canonicalize(G)
    assert abs(lam1 - lam0) <= 1e-12 * lam0
    for a, b in zip(can0, can1):
        assert np.allclose(a, b, rtol=0, atol=1e-13)
    masks = ws.butterfly_supports(n)
    for c, M in zip(can0, masks):
        assert np.all(c[M == 0] == 0)
        assert abs(np.linalg.norm(c) - 1.0) < 1e-14
    # every column of the first L-1 factors: equal norm and a positive diagonal
    for c in can0[:-1]:
        assert np.allclose(np.linalg.norm(c, axis=0), 1.0 / np.sqrt(n))
        assert np.all(np.diag(c) > 0)


@pytest.mark.parametrize("n", [16, 32, 64])
def test_butterfly_log_magnitudes_are_linear_of_rank_n_L_plus_1(n):
    # Every entry of a full butterfly product B_0 ... B_{L-1} is the product of
    # exactly one entry of each factor (a single path), so log|A| = M log|F|
    # for the 0/1 path-incidence matrix M (n^2 x 2nL, L ones per row), and
    # sign(A) is the parity of the path's signs.  M's rank is n (L + 1): the
    # 2nL factor values minus the n (L - 1) diagonal rescalings between
    # consecutive factors.  At n = 256 (L = 8) the map from log|A| to the
    # factors' log-magnitudes therefore has rank 2304, above the 512-wide
    # hidden layers of the configured 65536 -> 512 -> 512 -> 4096 network.
    from faust_amd import warmstart as ws
    rng = np.random.default_rng(n)
    masks = ws.butterfly_supports(n)
    L = len(masks)
    _, facs = ws.butterfly_sample(n, rng, 0.0)
    facs = [np.abs(f) + (m != 0) * 0.1 for f, m in zip(facs, masks)]  # strictly positive entries on the masks

    def logA(fs):
        P = np.eye(n)
        for f in fs:
            P = P @ f
        return np.log(P).ravel()

    v = ws.pack_values(facs, masks)
    base = logA(facs)
    M = np.empty((n * n, v.size))
    for p in range(v.size):  # d log A / d log F_p, exact: log A is linear in log F
        w = v.copy()
        w[p] *= np.e
        M[:, p] = logA(ws.unpack_values(w, masks)) - base
    assert np.allclose(M, np.round(M), atol=1e-9)
    M = np.round(M)
    assert np.all(M.sum(axis=1) == L)  # one entry of every factor on each path
    assert np.allclose(M @ np.log(v), base, atol=1e-9)
    assert np.linalg.matrix_rank(M) == n * (L + 1)
