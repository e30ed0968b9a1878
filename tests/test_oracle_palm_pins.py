"""Pins of the PALM4MSA restatement that do not go through the restatement's
own formulas:

* the reference's own known answer (test_FaustPy.py:1230-1263, :1380-1410):
  splin 5 + normcol 1.0 on the 500 x 32 matrix of config_compared_palm2.mat,
  200 iterations, relative error 0.270954109668 to 4 places, for both backends;
* normcol / normlin against their closed form (every column / row has 2-norm
  s and keeps its direction);
* update_lambda's sweep-order and aliasing rules, reconstructed from factors
  recorded at the previous iteration (faust_palm4msa2020.hpp:442-493 with
  TransformHelper::update, faust_TransformHelper.hpp:848-868);
* the Lipschitz step c = 1.001 lambda^2 ||L||_2^2 ||R||_2^2 of update_fact
  (faust_palm4msa2020.hpp:616) and the grad2 update
  D - (lambda/c) L^T (lambda L D R - A) R^T, rebuilt with numpy SVD norms
  and numpy products from the factors of a one-iteration run.
"""
import os

import numpy as np
import pytest

import oracle

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _relerr(r, M):
    P = np.eye(r.factors[0].shape[0])
    for f in r.factors:
        P = P @ f
    return np.linalg.norm(r.lam * P - M) / np.linalg.norm(M)


@pytest.mark.parametrize("backend", [2016, 2020])
def test_reference_palm4msa_known_answer(backend):
    d = np.load(os.path.join(G, "compared_palm2.npz"))
    M = d["data"]
    cons = [("splin", 5, 500, 32, True), ("normcol", 1.0, 32, 32, False)]
    run = oracle.palm4msa2016 if backend == 2016 else oracle.palm4msa
    r = run(M, cons, num_its=200, lambda_init=1.0, is_update_way_R2L=False)
    err = _relerr(r, M)
    # assertAlmostEqual(..., places=4): round(err - expected, 4) == 0
    assert round(err - float(d["expected_relerr"]), int(d["places"])) == 0
    # measured: 0.2709540369 for both backends (7e-8 from the reference's digits)
    assert abs(err - float(d["expected_relerr"])) < 1e-6
    # the factors satisfy their constraints
    assert (np.count_nonzero(r.factors[0], axis=1) <= 5).all()
    assert np.allclose(np.linalg.norm(r.factors[1], axis=0), 1.0, rtol=1e-14)


def test_prox_normcol_normlin_closed_form():
    rng = np.random.default_rng(3)
    M = rng.standard_normal((9, 6))
    M[:, 2] = 0.0
    for s in (1.0, 2.5):
        out = oracle.prox("normcol", M, s, normalized=False)
        nrm = np.linalg.norm(M, axis=0)
        for j in range(6):
            if nrm[j] == 0:
                assert not out[:, j].any()
            else:
                assert np.allclose(out[:, j], M[:, j] * (s / nrm[j]), rtol=1e-15, atol=0)
                assert abs(np.linalg.norm(out[:, j]) - s) < 4e-16 * s * 9
        outl = oracle.prox("normlin", M.T.copy(), s, normalized=False)
        assert np.array_equal(outl, out.T)
    assert not oracle.prox("normcol", M, 0.0, normalized=False).any()
    with pytest.raises(RuntimeError):
        oracle.prox("normcol", M, -1.0, normalized=False)
    # normalized: then divided by the Frobenius norm
    out = oracle.prox("normcol", M, 2.0, normalized=True)
    assert abs(np.linalg.norm(out) - 1.0) < 1e-15


def _lam_of(A, S):
    return np.trace(A.T @ S) / np.linalg.norm(S) ** 2


@pytest.mark.parametrize("fmt,packing,fresh", [
    ("dense", True, True),      # pL[1] = {S0}: pack_factors keeps a single dense factor, S.update writes into it
    ("dense", False, True),     # unpacked: pL[1] holds S0 itself
    ("sparse", False, True),    # unpacked: pL[1] holds S0 itself, S.update writes into it
    ("sparse", True, False),    # pack_factors turns the single sparse S0 into a new dense copy
    ("dynamic", True, False),   # S.replace: pL[1] keeps the object of the start of the sweep
    ("dynamic", False, False),
])
def test_update_lambda_right_to_left_two_factors(fmt, packing, fresh):
    # Two square factors updated right to left: update_lambda multiplies
    # pL[1] (built at the start of the sweep from S0) by the new S1.  Whether
    # pL[1] shows the new S0 depends on the factor format and on packing.
    rng = np.random.default_rng(11)
    A = rng.standard_normal((16, 16))
    cons = [("splin", 6, 16, 16), ("spcol", 6, 16, 16)]
    kw = dict(is_update_way_R2L=True, packing_RL=packing, factor_format=fmt)
    k = 9
    rk = oracle.palm4msa(A, cons, num_its=k, **kw)
    rprev = oracle.palm4msa(A, cons, num_its=k - 1, **kw)
    S0_new, S1_new = rk.factors
    S0_old = rprev.factors[0]
    lam_fresh = _lam_of(A, S0_new @ S1_new)
    lam_stale = _lam_of(A, S0_old @ S1_new)
    assert abs(lam_fresh - lam_stale) > 1e-8 * abs(lam_fresh)  # the two rules differ on this input
    want = lam_fresh if fresh else lam_stale
    assert abs(rk.lam - want) < 1e-12 * abs(want)


@pytest.mark.parametrize("fmt", ["dense", "sparse", "dynamic"])
@pytest.mark.parametrize("packing", [True, False])
def test_update_lambda_left_to_right_uses_new_factors(fmt, packing):
    # left to right, square factors: the last left product is rebuilt after
    # every update, so lambda comes from the final factors in every format
    rng = np.random.default_rng(12)
    A = rng.standard_normal((16, 16))
    cons = [("splin", 6, 16, 16), ("sp", 80, 16, 16), ("spcol", 6, 16, 16)]
    r = oracle.palm4msa(A, cons, num_its=7, is_update_way_R2L=False, packing_RL=packing, factor_format=fmt)
    S = r.factors[0] @ r.factors[1] @ r.factors[2]
    assert abs(r.lam - _lam_of(A, S)) < 1e-12 * abs(r.lam)


def _norm2(M):
    return np.linalg.svd(M, compute_uv=False)[0] if M.size else 0.0


def _rebuild_sweep(A, names, got, order, mult):
    """Factors of one sweep rebuilt from the recorded ones: factor f from
    prox(D - (1/c) L^T (L D R - A) R^T), lambda = 1, c = mult ||L||^2 ||R||^2."""
    I = np.eye(A.shape[0])
    cur = [I.copy() for _ in got]
    cur[order[0]] = np.zeros_like(I)
    out = [None] * len(got)
    for f in order:
        L = I.copy()
        for t in range(f):
            L = L @ cur[t]
        R = I.copy()
        for t in range(f + 1, len(got)):
            R = R @ cur[t]
        nL = _norm2(L) if f > 0 else 1.0
        nR = _norm2(R) if f < len(got) - 1 else 1.0
        c = mult * nR ** 2 * nL ** 2
        D = cur[f] - (1.0 / c) * (L.T @ (L @ cur[f] @ R - A) @ R.T)
        out[f] = oracle.prox(names[f][0], D, names[f][1])
        cur[f] = got[f]
    return out


@pytest.mark.parametrize("R2L", [False, True])
def test_lipschitz_step_and_grad2_one_sweep(R2L):
    # One iteration from the default start (the first updated factor 0, the
    # others I), lambda fixed at 1, accurate power iterations: each factor
    # after the sweep must equal the prox of the gradient step built with
    # numpy products and numpy SVD norms, c = 1.001 ||L||^2 ||R||^2.
    rng = np.random.default_rng(13)
    A = rng.standard_normal((12, 12))
    names = [("sp", 60), ("sp", 70), ("sp", 50)]
    cons = [(nm, k, 12, 12) for nm, k in names]
    r = oracle.palm4msa(A, cons, num_its=1, is_update_way_R2L=R2L, packing_RL=False, no_lambda=True,
                        norm2_max_iter=100000, norm2_threshold=1e-15)
    got = r.factors
    order = [2, 1, 0] if R2L else [0, 1, 2]
    want = _rebuild_sweep(A, names, got, order, 1.001)
    for f in order:
        assert np.array_equal(want[f] != 0, got[f] != 0), f"support of factor {f}"
        assert np.abs(want[f] - got[f]).max() < 1e-10 * np.abs(want[f]).max(), f"factor {f}"
    # a plain 1.0 multiplier (or a dropped square) would not match: the check sees c
    for mult in (1.0, 1.001 ** 2):
        other = _rebuild_sweep(A, names, got, order, mult)
        assert max(np.abs(other[f] - got[f]).max() for f in order[1:]) > 1e-6
