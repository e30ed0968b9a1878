"""Projectors, GEMM and reductions on the GPU vs the oracle (bit-exact)."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from faust_amd import _lib, proj

pytestmark = pytest.mark.gpu


def _tie_matrix(rng, m, n, levels=3):
    return np.asfortranarray(rng.integers(-levels, levels + 1, size=(m, n)).astype(np.float64))


@pytest.mark.parametrize("name", ["sp", "spcol", "splin", "splincol"])
@pytest.mark.parametrize("ties", [False, True])
@pytest.mark.parametrize("shape,k", [((32, 32), 2), ((64, 48), 5), ((17, 9), 1), ((40, 40), 39)])
def test_prox_bitexact(gpu, name, ties, shape, k):
    rng = np.random.default_rng(11)
    M = _tie_matrix(rng, *shape) if ties else np.asfortranarray(rng.standard_normal(shape))
    kk = k * 7 if name == "sp" else k
    P = getattr(proj, name)(shape, kk)
    got = P(M)
    ref = oracle.prox(name, M, kk)
    assert np.array_equal(got != 0, ref != 0), "support differs"
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("name", ["spcol", "splin", "splincol"])
def test_prox_large_k(gpu, name):
    rng = np.random.default_rng(2)
    M = _tie_matrix(rng, 1024, 1024, 2)
    got = getattr(proj, name)((1024, 1024), 512)(M)
    assert np.array_equal(got, oracle.prox(name, M, 512))


def test_prox_long_lines_global_scratch(gpu):
    # lines of 8192 with k = 3000 do not fit in LDS: per-line global scratch
    rng = np.random.default_rng(8)
    M = _tie_matrix(rng, 8192, 6, 3)
    got = proj.spcol((8192, 6), 3000)(M)
    assert np.array_equal(got, oracle.prox("spcol", M, 3000))


def test_prox_edge_k_and_pos(gpu):
    rng = np.random.default_rng(4)
    M = np.asfortranarray(rng.standard_normal((12, 10)))
    for k in (10, 12, 200):
        assert np.array_equal(proj.splin((12, 10), k)(M), oracle.prox("splin", M, k))
        assert np.array_equal(proj.spcol((12, 10), k)(M), oracle.prox("spcol", M, k))
    assert np.array_equal(proj.sp((12, 10), 30, pos=True)(M), oracle.prox("sp", M, 30, pos=True))
    S = (rng.random((12, 10)) < 0.3).astype(float)
    assert np.array_equal(proj.supp(S)(M), oracle.prox("supp", M, supp=S))
    with pytest.raises(_lib.FaustHipError):
        proj.sp((12, 10), 0)(M)  # zero matrix cannot be normalised (MatDense::normalize throws)


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("mnk", [(64, 64, 64), (70, 33, 129), (1, 200, 3)])
def test_gemm_bitexact(gpu, opA, opB, mnk):
    m, n, k = mnk
    rng = np.random.default_rng(9)
    A = np.asfortranarray(rng.standard_normal((m, k) if opA == "N" else (k, m)))
    B = np.asfortranarray(rng.standard_normal((k, n) if opB == "N" else (n, k)))
    C0 = np.asfortranarray(rng.standard_normal((m, n)))
    for alpha, beta in ((1.0, 0.0), (0.37, -1.0)):
        dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
        dB = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()
        dC = torch.from_numpy(np.ascontiguousarray(C0.T)).cuda()
        a, b = C.c_double(alpha), C.c_double(beta)
        _lib.check(gpu.fh_gemm(_lib.FH_F64, opA.encode(), opB.encode(), m, n, k, C.byref(a),
                               C.c_void_p(dA.data_ptr()), A.shape[0], C.c_void_p(dB.data_ptr()), B.shape[0],
                               C.byref(b), C.c_void_p(dC.data_ptr()), m, None))
        torch.cuda.synchronize()
        got = dC.cpu().numpy().T
        ref = oracle.gemm(A, B, opA, opB, alpha, beta, C0)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("n", [1, 255, 4096, 4097, 1 << 20])
def test_reductions_bitexact(gpu, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    dx, dy = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    out = C.c_double()
    _lib.check(gpu.fh_dot(_lib.FH_F64, C.c_void_p(dx.data_ptr()), C.c_void_p(dy.data_ptr()), n, C.byref(out), None))
    assert out.value == oracle.reduce_dot(x, y)


@pytest.mark.parametrize("name,shape,k", [("sp", (512, 512), 2048), ("sp", (300, 200), 59999),
                                          ("splincol", (1024, 1024), 512), ("spcol", (2000, 64), 700),
                                          ("splin", (96, 3000), 1)])
def test_prox_radix_threshold_paths(gpu, name, shape, k):
    # no ties (radix threshold decides), partial ties inside the top k (still
    # decided), tie groups straddling k (heap fallback), and zeros below k
    rng = np.random.default_rng(31)
    base = rng.standard_normal(shape)
    cases = [base,
             np.where(rng.random(shape) < 0.3, np.round(base, 1), base),
             np.round(base, 1),
             base * (rng.random(shape) < 0.002)]
    for M in cases:
        M = np.asfortranarray(M)
        if not M.any():
            continue
        got = getattr(proj, name)(shape, k)(M)
        ref = oracle.prox(name, M, k)
        assert np.array_equal(got != 0, ref != 0), "support differs"
        assert np.array_equal(got, ref)
