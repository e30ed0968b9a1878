"""The butterfly factorization on the GPU (fh_butterfly) vs the oracle's
restatement of faust_butterfly.hpp.

Singular vectors are defined up to sign, so factors are compared through
sign-free quantities: the entries' absolute values and the product.  A best
rank-1 approximation moves by about eps / (relative gap of the block's two
top singular values); on seeded Gaussian inputs that leaves differences of
1e-13 on products, and the bars below are
  * product: relative Frobenius difference <= 1e-11;
  * |factor|: max absolute difference <= 1e-9 * max |factor|.
Exact butterflies (products of factors on the butterfly supports) are
recovered to <= 1e-12 relative error at every order, and the reference's own
small / large expectations (misc/test/src/Python/test_butterfly.py:43-95)
hold on the GPU too.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from oracle import butterfly as ob
import faust_amd
from faust_amd import _lib, fact
from faust_amd import butterfly as hb

pytestmark = pytest.mark.gpu

_DIR = {"left": ob.LEFT, "right": ob.RIGHT, "bbtree": ob.BALANCED}
PROD_TOL = 1e-11
FAC_TOL = 1e-9


def _host(F):
    return [F.factors(i) for i in range(F.numfactors())]


def _dense(f):
    return f.toarray() if sp.issparse(f) else np.asarray(f)


def _compare(F, ref):
    got = [_dense(f) for f in _host(F)]
    exp = [_dense(f) for f in ref]
    assert len(got) == len(exp)
    pg, pe = ob.product(got), ob.product(exp)
    assert np.linalg.norm(pg - pe) <= PROD_TOL * np.linalg.norm(pe)
    for g, e in zip(got, exp):
        assert np.abs(np.abs(g) - np.abs(e)).max() <= FAC_TOL * np.abs(e).max()


@pytest.mark.parametrize("type", ["bbtree", "right", "left"])
@pytest.mark.parametrize("n", [4, 8, 16, 64, 256, 1024])
def test_random_vs_oracle(gpu, type, n):
    print(f"butterfly {type} n={n}", flush=True)
    rng = np.random.default_rng(100 + n)
    M = rng.standard_normal((n, n))
    F = fact.butterfly(M, type=type)
    _compare(F, ob.butterfly(M, _DIR[type]))
    L = int(np.log2(n))
    kinds = [f.info()[0] for f in F._facs]
    want = _lib.FH_MAT_CSR if type == "bbtree" else _lib.FH_MAT_DENSE
    assert kinds == [want] * L
    # every factor on its butterfly support
    for i, f in enumerate(_host(F)):
        d = _dense(f)
        assert not np.any(d[hb.support_csr(n, i).toarray() == 0])


@pytest.mark.parametrize("type", ["bbtree", "right", "left"])
@pytest.mark.parametrize("n", [32, 512, 4096])
def test_exact_butterfly_recovered(gpu, type, n):
    print(f"butterfly exact {type} n={n}", flush=True)
    rng = np.random.default_rng(n)
    B = faust_amd.Faust(hb.rand_butterfly_factors(n, rng=rng))
    M = B.toarray()
    F = fact.butterfly(M, type=type)
    assert np.linalg.norm(F.toarray() - M) <= 1e-12 * np.linalg.norm(M)


def test_hadamard_recovered(gpu):
    for n in (8, 1024):
        H = oracle.hadamard(n)
        for type in ("bbtree", "right", "left"):
            F = fact.butterfly(H, type=type)
            assert np.linalg.norm(F.toarray() - H) <= 1e-12 * np.linalg.norm(H), (n, type)


@pytest.mark.parametrize("pname", ["bitrev", "rand"])
@pytest.mark.parametrize("hname", ["rand_butterfly", "hadamard"])
def test_reference_expectations(gpu, pname, hname):
    """test_butterfly.py:43-95 through the GPU path (bit reversal is its own
    transpose: cases 5 and 7 are cases 4 and 8 again)."""
    from tests.test_butterfly import _cases, _expected, _rand_bfly_dense, _rand_perm
    rng = np.random.default_rng(7)
    N = 32
    P = hb.bitrev_perm(N) if pname == "bitrev" else _rand_perm(N, rng)
    H = _rand_bfly_dense(N, rng) if hname == "rand_butterfly" else oracle.hadamard(N)
    for i, (mat, p) in enumerate(_cases(H, P)):
        F = fact.butterfly(np.asarray(mat), type="bbtree", perm=p)
        e = np.linalg.norm(F.toarray() - mat) / np.linalg.norm(mat)
        assert ("small" if e < 1e-12 else "large") == _expected(i, P), (i, e)


@pytest.mark.parametrize("type", ["bbtree", "right", "left"])
@pytest.mark.parametrize("mul_perm", [True, False])
def test_perm_vs_oracle(gpu, type, mul_perm):
    rng = np.random.default_rng(21)
    n = 64
    M = rng.standard_normal((n, n))
    p = rng.permutation(n)
    F = fact.butterfly(M, type=type, perm=p, mul_perm=mul_perm)
    ref = ob.butterfly(M, _DIR[type], p, mul_perm=mul_perm)
    assert F.numfactors() == (6 if mul_perm else 7)
    _compare(F, ref)
    if mul_perm:
        assert F._facs[-1].info()[0] == _lib.FH_MAT_CSR


@pytest.mark.parametrize("perm", [False, True])
def test_leaf_factor_views(gpu, perm):
    """The CSR leaf factors are built on the device with their CSC twin and
    the ELL views of both: rows sorted, 2 entries per row, and F @ X, F.T @ X
    (rows of the twin) agree with the product of the host copies."""
    rng = np.random.default_rng(31)
    n = 256
    M = rng.standard_normal((n, n))
    p = rng.permutation(n) if perm else None
    F = fact.butterfly(M, type="bbtree", perm=p)
    facs = _host(F)
    for k, f in enumerate(facs):
        assert F._facs[k].info()[0] == _lib.FH_MAT_CSR
        S = F._facs[k].to_host()
        assert sp.issparse(S) and S.nnz == 2 * n and S.has_sorted_indices
        assert np.array_equal(np.diff(S.indptr), np.full(n, 2))
        assert np.array_equal(S.toarray(), f)
        assert F._facs[k].info()[4] == 2 * n
    D = ob.product(facs)
    X = rng.standard_normal((n, 5))
    for got, ref in ((F @ X, D @ X), (F.T @ X, D.T @ X)):
        assert np.linalg.norm(np.asarray(got) - ref) <= 1e-12 * np.linalg.norm(ref)


def test_dense_leaf_nnz(gpu):
    """Dense leaf factors (left / right) report their nonzero count without
    scanning the n x n array: it equals the count of the dense copy."""
    rng = np.random.default_rng(8)
    n = 64
    M = rng.standard_normal((n, n))
    M[:, 3] = 0.0  # exact zeros in some node values
    for type in ("right", "left"):
        F = fact.butterfly(M, type=type)
        for k, f in enumerate(_host(F)):
            assert F._facs[k].info()[4] == np.count_nonzero(_dense(f)), (type, k)


@pytest.mark.parametrize("n", [64, 1024, 4096])
def test_valu_gram_kernel(gpu, n, monkeypatch):
    """FH_BFLY_MFMA=0 runs the VALU Gram kernel at every block size: same
    bars against the oracle, and the same product as the MFMA kernel within
    the product bar."""
    rng = np.random.default_rng(300 + n)
    M = rng.standard_normal((n, n))
    Fm = fact.butterfly(M, type="bbtree")
    monkeypatch.setenv("FH_BFLY_MFMA", "0")
    Fv = fact.butterfly(M, type="bbtree")
    monkeypatch.delenv("FH_BFLY_MFMA")
    if n <= 1024:
        _compare(Fv, ob.butterfly(M, ob.BALANCED))
    pm, pv = Fm.toarray(), Fv.toarray()
    assert np.linalg.norm(pm - pv) <= PROD_TOL * np.linalg.norm(pm)


@pytest.mark.parametrize("type", ["right", "left"])
@pytest.mark.parametrize("n", [4096, 8192])
def test_long_d2_blocks(gpu, type, n, monkeypatch):
    """Blocks with a side of 2 and a long side of 2048..8192 (the first splits
    of right / left at these orders) run one workgroup per block with the
    block in registers: exact butterflies are recovered to 1e-12, and a
    Gaussian matrix's factorization has the same product as with the
    one-wave kernel (FH_BFLY_D2_WIDE=0), within the product bar."""
    from faust_amd.butterfly import rand_butterfly_factors
    rng = np.random.default_rng(n)
    print(f"long d2 blocks {type} n={n}", flush=True)
    B = faust_amd.Faust(rand_butterfly_factors(n, rng=rng)).toarray()
    F = fact.butterfly(B, type=type)
    assert np.linalg.norm(F.toarray() - B) <= 1e-12 * np.linalg.norm(B)
    M = rng.standard_normal((n, n))
    Fw = fact.butterfly(M, type=type)
    monkeypatch.setenv("FH_BFLY_D2_WIDE", "0")
    Fn = fact.butterfly(M, type=type)
    monkeypatch.delenv("FH_BFLY_D2_WIDE")
    pw, pn = Fw.toarray(), Fn.toarray()
    assert np.linalg.norm(pw - pn) <= PROD_TOL * np.linalg.norm(pn)


def test_named_perms(gpu):
    """'bitrev' = its index vector; 'default_8' keeps the best of the eight."""
    rng = np.random.default_rng(5)
    n = 16
    M = rng.standard_normal((n, n))
    F1 = fact.butterfly(M, perm="bitrev")
    F2 = fact.butterfly(M, perm=hb.perm2indices(hb.bitrev_perm(n)))
    assert np.array_equal(F1.toarray(), F2.toarray())
    F8 = fact.butterfly(M, perm="default_8")
    errs = [ob.rel_err(ob.butterfly(M, ob.BALANCED, p), M) for p in hb.default_8_perms(n)]
    e8 = np.linalg.norm(F8.toarray() - M) / np.linalg.norm(M)
    assert abs(e8 - min(errs)) <= 1e-10


def test_identity_perm_same_as_none(gpu):
    rng = np.random.default_rng(3)
    M = rng.random((4, 4))
    F1 = fact.butterfly(M, type="bbtree")
    F2 = fact.butterfly(M, type="bbtree", perm=np.arange(4))
    assert np.array_equal(F1.toarray(), F2.toarray())


@pytest.mark.parametrize("type", ["bbtree", "right", "left"])
def test_float32_input(gpu, type):
    rng = np.random.default_rng(9)
    M = rng.standard_normal((64, 64)).astype(np.float32)
    F = fact.butterfly(M, type=type)
    assert F.dtype == np.float32
    ref = ob.product(ob.butterfly(M.astype(np.float64), _DIR[type]))
    # FP32 factors and an FP32 product: a few float32 roundings per factor
    assert np.linalg.norm(F.toarray() - ref) <= 1e-5 * np.linalg.norm(ref)


def test_errors(gpu):
    rng = np.random.default_rng(0)
    with pytest.raises(ValueError):
        fact.butterfly(rng.standard_normal((8, 4)))
    with pytest.raises(ValueError):
        fact.butterfly(rng.standard_normal((12, 12)))
    with pytest.raises(ValueError):
        fact.butterfly(rng.standard_normal((8, 8)), type="up")
    with pytest.raises(ValueError):
        fact.butterfly(rng.standard_normal((8, 8)), perm=[0, 1, 2])
    with pytest.raises(ValueError):
        fact.butterfly(rng.standard_normal((8, 8)), perm=[0, 1, 2, 3, 4, 5, 6, 8])
    with pytest.raises(ValueError):
        fact.butterfly(rng.standard_normal((8, 8)), diag_opt=True, mul_perm=True)
    with pytest.raises(NotImplementedError):
        fact.butterfly(rng.standard_normal((8, 8)), diag_opt=True)
    with pytest.raises(_lib.FaustHipError):
        fact.butterfly(rng.standard_normal((2, 2)))
    with pytest.raises(_lib.FaustHipError):  # repeated index: not a permutation
        fact.butterfly(rng.standard_normal((8, 8)), perm=[0, 0, 2, 3, 4, 5, 6, 7])


def test_rand_butterfly(gpu):
    np.random.seed(4)
    F = faust_amd.rand_butterfly(32)
    assert F.numfactors() == 5
    for i, f in enumerate(_host(F)):
        assert np.array_equal(_dense(f) != 0, hb.support_csr(32, i).toarray() != 0)
    G = fact.butterfly(F.toarray())
    assert np.linalg.norm(G.toarray() - F.toarray()) <= 1e-12 * np.linalg.norm(F.toarray())


@pytest.mark.parametrize("type", ["bbtree", "right", "left"])
def test_device_input_same_bits(gpu, type):
    """A CUDA tensor (row-major, column-major, or a strided slice) gives the
    factors of the host path bit for bit: the root kernels only copy."""
    import torch
    rng = np.random.default_rng(44)
    n = 128
    M = np.asfortranarray(rng.standard_normal((n, n)))
    p = rng.permutation(n)
    ref = fact.butterfly(M, type=type, perm=p)
    want = [_dense(f) for f in _host(ref)]
    big = torch.zeros((n + 3, n + 5), dtype=torch.float64, device="cuda")
    big[:n, :n] = torch.from_numpy(np.ascontiguousarray(M)).cuda()
    cm = torch.from_numpy(np.ascontiguousarray(M.T)).cuda().t()  # column-major view
    for T in (torch.from_numpy(np.ascontiguousarray(M)).cuda(), cm, big[:n, :n]):
        F = fact.butterfly(T, type=type, perm=p)
        got = [_dense(f) for f in _host(F)]
        assert len(got) == len(want)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)


def test_bbtree_order_32768(gpu):
    """The largest bbtree order: 256 x 128 blocks at the root (a 128 x 128
    Gram matrix per block).  An exact butterfly of order 32768 is formed on
    the GPU (the Faust applied to the identity), factorized in place, and the
    result checked on random columns."""
    import torch
    n = 32768
    rng = np.random.default_rng(32768)
    B = faust_amd.Faust(hb.rand_butterfly_factors(n, rng=rng))
    eye = torch.eye(n, dtype=torch.float64, device="cuda")
    out = B.matmul_device(eye)  # (n, n) row-major = B column-major
    del eye
    torch.cuda.empty_cache()
    F = fact.butterfly(out.t(), type="bbtree")
    del out
    torch.cuda.empty_cache()
    assert F.numfactors() == 15
    X = rng.standard_normal((n, 4))
    Y, R = F @ X, B @ X
    assert np.linalg.norm(Y - R) <= 1e-12 * np.linalg.norm(R)
