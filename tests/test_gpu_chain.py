"""Faust x dense on the GPU vs the oracle: bit-exact in FP64 (canonical order)."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import faust_amd

pytestmark = pytest.mark.gpu


def _cfg1_factors(seed=0xFA057, n=1024, nnz=8, J=5):
    # 5 CSR factors 1024x1024, exactly 8 distinct uniform columns per row, N(0,1)
    return [faust_amd.rand_csr(n, n, nnz, np.random.default_rng(seed + f)) for f in range(J)]


@pytest.mark.parametrize("op", ["N", "T"])
def test_cfg1_chain_bitexact(gpu, op):
    facs = _cfg1_factors()
    X = np.asfortranarray(np.random.default_rng(0xFA057 + 100).standard_normal((1024, 64)))
    F = faust_amd.Faust(facs)
    Y = (F @ X) if op == "N" else (F.T @ X)
    ref = oracle.chain_multiply(facs, X, op)
    assert Y.shape == ref.shape
    assert np.array_equal(Y, ref)


@pytest.mark.parametrize("shapes", [[(7, 13), (13, 5), (5, 9)], [(64, 1), (1, 64)], [(33, 33)]])
@pytest.mark.parametrize("kinds", ["sparse", "dense", "mixed"])
def test_ragged_chains(gpu, shapes, kinds):
    rng = np.random.default_rng(3)
    facs = []
    for i, (r, c) in enumerate(shapes):
        dense = kinds == "dense" or (kinds == "mixed" and i % 2 == 1)
        a = rng.standard_normal((r, c)) * (rng.random((r, c)) < 0.4)
        facs.append(np.asfortranarray(a) if dense else sp.csr_matrix(a))
    F = faust_amd.Faust(facs)
    for p in (1, 3, 17):
        X = np.asfortranarray(rng.standard_normal((shapes[-1][1], p)))
        assert np.array_equal(F @ X, oracle.chain_multiply(facs, X, "N"))
        X2 = np.asfortranarray(rng.standard_normal((shapes[0][0], p)))
        assert np.array_equal(F.T @ X2, oracle.chain_multiply(facs, X2, "T"))


def test_empty_rows_and_vector(gpu):
    rng = np.random.default_rng(5)
    a = sp.random(50, 40, density=0.05, random_state=1, format="csr")  # many empty rows
    b = sp.random(40, 60, density=0.1, random_state=2, format="csr")
    F = faust_amd.Faust([a, b])
    x = rng.standard_normal(60)
    y = F @ x
    assert y.shape == (50,)
    assert np.array_equal(y, oracle.chain_multiply([a, b], x.reshape(-1, 1), "N").ravel())


def test_get_product_and_norm2(gpu):
    facs = _cfg1_factors(n=256, nnz=4, J=4)
    F = faust_amd.Faust(facs)
    assert np.array_equal(F.toarray(), oracle.get_product(facs))
    nrm, _ = oracle.spectral_norm(facs)
    assert F.norm(2) == nrm


def test_wht_faust(gpu):
    F = faust_amd.wht(256)
    H = oracle.hadamard(256)
    assert np.abs(F.toarray() - H).max() < 1e-14
    assert abs(F.norm(2) - 1.0) < 1e-6


@pytest.fixture(params=["0", "1"], ids=["persistent", "wide"])
def power_form(request, monkeypatch):
    # FH_POWER_WIDE forces the one-workgroup or the one-launch-per-operator form
    monkeypatch.setenv("FH_POWER_WIDE", request.param)
    return request.param


@pytest.mark.parametrize("case", ["sparse256", "dense_rect", "mixed", "long_rows1024"])
def test_norm2_both_forms_bitexact(gpu, power_form, case):
    rng = np.random.default_rng(17)
    if case == "sparse256":
        facs = _cfg1_factors(n=256, nnz=4, J=4)
    elif case == "dense_rect":
        facs = [np.asfortranarray(rng.standard_normal((300, 200))), np.asfortranarray(rng.standard_normal((200, 310)))]
    elif case == "mixed":
        facs = [sp.csr_matrix(rng.standard_normal((130, 70)) * (rng.random((130, 70)) < 0.2)),
                np.asfortranarray(rng.standard_normal((70, 90))),
                faust_amd.rand_csr(90, 90, 5, rng)]
    else:  # residual-like factor: 1024 x 1024 with 600 nnz per row, plus a dense one
        facs = [faust_amd.rand_csr(1024, 1024, 600, rng), np.asfortranarray(rng.standard_normal((1024, 1024)))]
    F = faust_amd.Faust(facs)
    nrm, _ = oracle.spectral_norm(facs)
    assert F.norm(2) == nrm
