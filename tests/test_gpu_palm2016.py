"""The 2016 PALM4MSA / hierarchical backend on the GPU vs the oracle's
restatement of Faust::Palm4MSA and Faust::HierarchicalFact: factors, supports,
lambda and the per-iteration lambda, bit for bit."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import faust_amd
from faust_amd import fact, factparams

pytestmark = pytest.mark.gpu


def _check_same(res, ref):
    assert res.lam == ref.lam
    assert len(res.factors) == len(ref.factors)
    for a, b in zip(res.dense_factors(), ref.factors):
        assert np.array_equal(a != 0, b != 0), "support differs"
        assert np.array_equal(a, b)
    assert np.array_equal(res.traj, ref.traj, equal_nan=True)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("R2L", [False, True])
def test_palm4msa2016_bitexact(gpu, mode, R2L):
    rng = np.random.default_rng(31)
    A = np.asfortranarray(rng.standard_normal((32, 24)))
    cons = [factparams.ConstraintInt("splin", 32, 16, 4), factparams.ConstraintInt("sp", 16, 16, 64),
            factparams.ConstraintInt("spcol", 16, 24, 3)]
    p = factparams.ParamsPalm4MSA(cons, factparams.StoppingCriterion(num_its=12), is_update_way_R2L=R2L,
                                  grad_calc_opt_mode=mode)
    res = fact.palm4msa(A, p, backend=2016, ret_raw=True)
    ref = oracle.palm4msa2016(A, [("splin", 4, 32, 16), ("sp", 64, 16, 16), ("spcol", 3, 16, 24)], num_its=12,
                              record_traj=True, is_update_way_R2L=R2L, grad_calc_opt_mode=mode)
    _check_same(res, ref)


def test_palm4msa2016_init_facts_and_faust(gpu):
    rng = np.random.default_rng(32)
    A = np.asfortranarray(rng.standard_normal((20, 20)))
    init = [np.asfortranarray(rng.standard_normal((20, 20))), np.asfortranarray(np.eye(20))]
    cons = [factparams.ConstraintInt("splincol", 20, 20, 5), factparams.ConstraintInt("splincol", 20, 20, 5)]
    p = factparams.ParamsPalm4MSA(cons, factparams.StoppingCriterion(num_its=8), init_facts=init)
    res = fact.palm4msa(A, p, backend=2016, ret_raw=True)
    ref = oracle.palm4msa2016(A, [("splincol", 5, 20, 20), ("splincol", 5, 20, 20)], num_its=8, init_facts=init,
                              record_traj=True)
    _check_same(res, ref)
    # the returned Faust: dense unless 2*nnz + rows + 1 < rows*cols, lambda in
    # the factor of fewest bytes
    F, lam = fact.palm4msa(A, p, backend=2016, ret_lambda=True)
    assert lam == ref.lam
    P = ref.lam * (ref.factors[0] @ ref.factors[1])
    assert np.linalg.norm(F.toarray() - P) <= 1e-13 * np.linalg.norm(P)


@pytest.mark.parametrize("R2L,left", [(True, False), (False, False), (True, True)])
def test_hierarchical2016_hadamard32_bitexact(gpu, R2L, left):
    # the reference's config_HADAMARD_32.mat configuration (5 factors,
    # splincol 2 / 16-8-4-2, 30 + 30 iterations), both update ways and both
    # factor sides
    with open(os.path.join(os.path.dirname(__file__), "golden", "hadamard32_config.json")) as f:
        cfg = json.load(f)
    H = oracle.hadamard(32)
    fc = [factparams.ConstraintInt(c[0], c[2], c[3], c[1]) for c in cfg["fact_constraints"]]
    rc = [factparams.ConstraintInt(c[0], c[2], c[3], c[1]) for c in cfg["res_constraints"]]
    p = factparams.ParamsHierarchical(fc, rc, factparams.StoppingCriterion(num_its=cfg["niter1"]),
                                      factparams.StoppingCriterion(num_its=cfg["niter2"]), is_update_way_R2L=R2L,
                                      is_fact_side_left=left)
    res = fact.hierarchical(H, p, backend=2016, ret_raw=True)
    ref = oracle.hierarchical2016(H, [tuple(c) for c in cfg["fact_constraints"]],
                                  [tuple(c) for c in cfg["res_constraints"]], num_its_local=cfg["niter1"],
                                  num_its_global=cfg["niter2"], record_traj=True, is_update_way_R2L=R2L,
                                  is_fact_side_left=left)
    _check_same(res, ref)
    if R2L and not left:
        assert abs(res.lam - cfg["expected_lambda"]) <= cfg["lambda_tolerance"]
        # the returned Faust: sparse factors, lambda in the first one
        F = fact.hierarchical(H, p, backend=2016)
        f0 = F.factors(0)
        assert np.array_equal(f0, ref.lam * ref.factors[0])
        assert np.linalg.norm(F.toarray() - H) <= 1e-10 * np.linalg.norm(H)


def test_2016_error_criterion_is_refused(gpu):
    A = np.eye(8)
    cons = [factparams.ConstraintInt("sp", 8, 8, 8), factparams.ConstraintInt("sp", 8, 8, 8)]
    p = factparams.ParamsPalm4MSA(cons, factparams.StoppingCriterion(tol=1e-3))
    with pytest.raises(faust_amd.FaustHipError):
        fact.palm4msa(A, p, backend=2016)
