"""PALM4MSA / hierarchical on the GPU vs the oracle: factors, supports, lambda
and the per-iteration relative-error trajectory, bit for bit."""
import numpy as np
import pytest

import oracle
from faust_amd import fact, factparams

pytestmark = pytest.mark.gpu


def _check_same(res, ref):
    assert res.lam == ref.lam
    assert len(res.factors) == len(ref.factors)
    for a, b in zip(res.dense_factors(), ref.factors):
        assert np.array_equal(a != 0, b != 0), "support differs"
        assert np.array_equal(a, b)
    assert res.kinds() == ref.kinds
    if len(ref.traj):
        assert np.array_equal(res.traj, ref.traj)


@pytest.mark.parametrize("n_log2,its", [(4, 10), (5, 30), (6, 20)])
def test_hierarchical_hadamard(gpu, n_log2, its):
    n = 2 ** n_log2
    H = oracle.hadamard(n)
    p = factparams.ParamsHierarchicalWHT(n_log2, num_its=its)
    res = fact.hierarchical(H, p, ret_raw=True)
    fac, resc, kw = oracle.wht_params(n_log2, its)
    ref = oracle.hierarchical(H, fac, resc, record_traj=True, **kw)
    _check_same(res, ref)
    P = np.eye(n)
    for f in ref.factors:
        P = P @ f
    assert np.linalg.norm(ref.lam * P - H) / np.linalg.norm(H) < 1e-10


@pytest.mark.parametrize("R2L", [False, True])
@pytest.mark.parametrize("packing", [True, False])
@pytest.mark.parametrize("fmt", ["dynamic", "dense", "sparse"])
def test_palm4msa_small(gpu, R2L, packing, fmt):
    rng = np.random.default_rng(21)
    A = np.asfortranarray(rng.standard_normal((32, 24)))
    cons = [factparams.ConstraintInt("splin", 32, 16, 4), factparams.ConstraintInt("sp", 16, 16, 64),
            factparams.ConstraintInt("spcol", 16, 24, 3)]
    p = factparams.ParamsPalm4MSA(cons, factparams.StoppingCriterion(num_its=15), is_update_way_R2L=R2L,
                                  packing_RL=packing, factor_format=fmt)
    res = fact.palm4msa(A, p, ret_raw=True)
    ref = oracle.palm4msa(A, [("splin", 4, 32, 16), ("sp", 64, 16, 16), ("spcol", 3, 16, 24)], num_its=15,
                          record_traj=True, is_update_way_R2L=R2L, packing_RL=packing, factor_format=fmt)
    _check_same(res, ref)


def test_hierarchical_hadamard_wide_power(gpu, monkeypatch):
    # the one-launch-per-operator power iteration gives the same trajectory
    monkeypatch.setenv("FH_POWER_WIDE", "1")
    n = 32
    H = oracle.hadamard(n)
    p = factparams.ParamsHierarchicalWHT(5, num_its=20)
    res = fact.hierarchical(H, p, ret_raw=True)
    fac, resc, kw = oracle.wht_params(5, 20)
    ref = oracle.hierarchical(H, fac, resc, record_traj=True, **kw)
    _check_same(res, ref)


@pytest.mark.parametrize("R2L", [False, True])
@pytest.mark.parametrize("fmt", ["dynamic", "dense"])
def test_palm4msa_grad1_bitexact(gpu, R2L, fmt):
    # the USE_GRAD1 gradient (the reference's GPU2 form) against the oracle's
    rng = np.random.default_rng(22)
    A = np.asfortranarray(rng.standard_normal((32, 24)))
    cons = [factparams.ConstraintInt("splin", 32, 16, 4), factparams.ConstraintInt("sp", 16, 16, 64),
            factparams.ConstraintInt("spcol", 16, 24, 3)]
    p = factparams.ParamsPalm4MSA(cons, factparams.StoppingCriterion(num_its=12), is_update_way_R2L=R2L,
                                  factor_format=fmt)
    p.use_grad1 = True
    res = fact.palm4msa(A, p, ret_raw=True)
    ref = oracle.palm4msa(A, [("splin", 4, 32, 16), ("sp", 64, 16, 16), ("spcol", 3, 16, 24)], num_its=12,
                          record_traj=True, is_update_way_R2L=R2L, factor_format=fmt, use_grad1=True)
    _check_same(res, ref)


def test_hierarchical_grad1_bitexact(gpu, monkeypatch):
    monkeypatch.setenv("USE_GRAD1", "1")
    H = oracle.hadamard(32)
    res = fact.hierarchical(H, factparams.ParamsHierarchicalWHT(5, num_its=15), ret_raw=True)
    fac, resc, kw = oracle.wht_params(5, 15)
    ref = oracle.hierarchical(H, fac, resc, record_traj=True, use_grad1=True, **kw)
    _check_same(res, ref)


def test_hierarchical_hadamard_256_bitexact(gpu):
    # 8 levels on the 256 Hadamard matrix (residuals k = 128 ... 2), 5
    # iterations per PALM4MSA: factors, lambda and every per-iteration error
    H = oracle.hadamard(256)
    res = fact.hierarchical(H, factparams.ParamsHierarchicalWHT(8, num_its=5), ret_raw=True)
    fac, resc, kw = oracle.wht_params(8, 5)
    ref = oracle.hierarchical(H, fac, resc, record_traj=True, **kw)
    _check_same(res, ref)
