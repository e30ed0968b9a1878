"""PALM4MSA on the GPU against the oracle and the reference's own known answer.

* The reference's PALM4MSA test (misc/test/src/Python/test_FaustPy.py:1230-1263
  for backend 2016, :1380-1410 for backend 2020): splin 5 + normcol 1.0 on the
  500 x 32 matrix of config_compared_palm2.mat, 200 iterations; relative error
  0.270954109668 to 4 places.  The GPU run equals the oracle bit for bit.
* The factor formats whose updates are written in place (AllDense /
  AllSparse, TransformHelper::update, faust_TransformHelper.hpp:848-868),
  which change what update_lambda sees in right-to-left sweeps.
* normcol / normlin (faust_prox.hpp:205-255) against the oracle.
"""
import os

import numpy as np
import pytest

import oracle
from faust_amd import fact, factparams, proj

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _same(res, ref):
    assert res.lam == ref.lam
    assert len(res.factors) == len(ref.factors)
    for a, b in zip(res.dense_factors(), ref.factors):
        assert np.array_equal(a != 0, b != 0), "support differs"
        assert np.array_equal(a, b)
    assert res.kinds() == ref.kinds
    if len(ref.traj):
        assert np.array_equal(res.traj[:, [0, 1, 3]], ref.traj[:, [0, 1, 3]])
        assert np.array_equal(np.isnan(res.traj[:, 2]), np.isnan(ref.traj[:, 2]))
        ok = ~np.isnan(ref.traj[:, 2])
        assert np.array_equal(res.traj[ok, 2], ref.traj[ok, 2])


@pytest.mark.parametrize("backend", [2016, 2020])
def test_reference_palm4msa_known_answer(gpu, backend):
    d = np.load(os.path.join(G, "compared_palm2.npz"))
    M = d["data"]
    print(f"reference PALM4MSA known answer, backend {backend}, 500x32, 200 its", flush=True)
    cons = [factparams.ConstraintInt("splin", 500, 32, 5),
            factparams.ConstraintReal("normcol", 32, 32, 1.0)]
    p = factparams.ParamsPalm4MSA(cons, factparams.StoppingCriterion(num_its=200), init_facts=None,
                                  is_update_way_R2L=False, init_lambda=1.0, constant_step_size=False)
    res = fact.palm4msa(M, p, backend=backend, ret_raw=True)
    ocons = [("splin", 5, 500, 32, True), ("normcol", 1.0, 32, 32, False)]
    run = oracle.palm4msa2016 if backend == 2016 else oracle.palm4msa
    ref = run(M, ocons, num_its=200, lambda_init=1.0, is_update_way_R2L=False, record_traj=True)
    _same(res, ref)
    # the reference's assertion, on the Faust that fact.palm4msa returns
    F, lam = fact.palm4msa(M, p, ret_lambda=True, backend=backend)
    E = F.toarray() - M
    err = np.linalg.norm(E) / np.linalg.norm(M)
    assert lam == res.lam
    assert round(err - float(d["expected_relerr"]), int(d["places"])) == 0
    assert abs(err - float(d["expected_relerr"])) < 1e-6


@pytest.mark.parametrize("fmt", ["dense", "sparse", "dynamic"])
@pytest.mark.parametrize("packing", [True, False])
@pytest.mark.parametrize("R2L", [True, False])
def test_palm4msa_two_factors_formats(gpu, fmt, packing, R2L):
    # two factors: update_lambda reads pL[1] / pR[0], which hold the factor
    # objects themselves unless packing copied them
    rng = np.random.default_rng(11)
    A = np.asfortranarray(rng.standard_normal((16, 16)))
    cons = [factparams.ConstraintInt("splin", 16, 16, 6), factparams.ConstraintInt("spcol", 16, 16, 6)]
    p = factparams.ParamsPalm4MSA(cons, factparams.StoppingCriterion(num_its=9), is_update_way_R2L=R2L,
                                  packing_RL=packing, factor_format=fmt)
    res = fact.palm4msa(A, p, ret_raw=True)
    ref = oracle.palm4msa(A, [("splin", 6, 16, 16), ("spcol", 6, 16, 16)], num_its=9, record_traj=True,
                          is_update_way_R2L=R2L, packing_RL=packing, factor_format=fmt)
    _same(res, ref)


@pytest.mark.parametrize("fmt", ["dense", "sparse"])
def test_hierarchical_hadamard_formats(gpu, fmt):
    # every local PALM has two factors updated right to left
    H = oracle.hadamard(32)
    p = factparams.ParamsHierarchicalWHT(5, num_its=20)
    p.factor_format = fmt
    res = fact.hierarchical(H, p, ret_raw=True)
    fac, resc, kw = oracle.wht_params(5, 20)
    ref = oracle.hierarchical(H, fac, resc, record_traj=True, factor_format=fmt, **kw)
    _same(res, ref)


@pytest.mark.parametrize("name", ["normcol", "normlin"])
def test_prox_normcol_normlin(gpu, name):
    rng = np.random.default_rng(5)
    for shape in ((9, 6), (300, 40), (40, 5000)):
        M = rng.standard_normal(shape)
        M[:, 1] = 0.0
        M[2, :] = 0.0
        for s, normalized, pos in ((1.0, False, False), (2.5, True, False), (0.7, False, True), (0.0, False, False)):
            P = (proj.normcol if name == "normcol" else proj.normlin)(shape, s, normalized=normalized, pos=pos)
            got = P(M)
            want = oracle.prox(name, M, s, normalized=normalized, pos=pos)
            assert np.array_equal(got, want), (shape, s, normalized, pos)


@pytest.mark.parametrize("rel_tol,use_grad1,its", [(0.9, False, 2), (0.7, False, 6), (0.62, True, 14),
                                                 (1e-9, False, 25)])
def test_palm4msa_error_stopping_criterion(gpu, rel_tol, use_grad1, its):
    # StoppingCriterion's error branch (faust_StoppingCriterion.hpp:113-150):
    # stop once the residual norm measured at the last factor update drops
    # below tol, else after maxiter iterations; tol relative to ||A||
    # (relerr=True).  The GPU stops at the same iteration as the oracle, with
    # the same bits.
    rng = np.random.default_rng(31)
    A = np.asfortranarray(rng.standard_normal((32, 24)))
    cons = [factparams.ConstraintInt("splin", 32, 16, 4), factparams.ConstraintInt("sp", 16, 16, 64),
            factparams.ConstraintInt("spcol", 16, 24, 3)]
    sc = factparams.StoppingCriterion(tol=rel_tol, maxiter=25, relerr=True, relmat=A)
    p = factparams.ParamsPalm4MSA(cons, sc)
    p.use_grad1 = use_grad1
    res = fact.palm4msa(A, p, ret_raw=True)
    ref = oracle.palm4msa(A, [("splin", 4, 32, 16), ("sp", 64, 16, 16), ("spcol", 3, 16, 24)],
                          tol=rel_tol * np.linalg.norm(A), maxiter=25, record_traj=True, use_grad1=use_grad1)
    _same(res, ref)
    assert res.traj.shape[0] == its
