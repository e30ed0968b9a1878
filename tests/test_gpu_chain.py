"""Faust x dense on the GPU vs the oracle: bit-exact in FP64 (canonical order)."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import faust_amd

pytestmark = pytest.mark.gpu


def _csr_chain_factors(seed=0xFA057, n=1024, nnz=8, J=5):
    # 5 CSR factors 1024x1024, exactly 8 distinct uniform columns per row, N(0,1)
    return [faust_amd.rand_csr(n, n, nnz, np.random.default_rng(seed + f)) for f in range(J)]


@pytest.mark.parametrize("op", ["N", "T"])
def test_csr1024x5_chain_bitexact(gpu, op):
    facs = _csr_chain_factors()
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
    facs = _csr_chain_factors(n=256, nnz=4, J=4)
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


def _short_factor(m, n, rng, k, wide_row=False):
    """m x n CSR whose rows and columns have at most k entries (k random partial
    matchings of rows to columns, each keeping a random half of its pairs), so
    that the factor and its transpose both have short rows; wide_row: one row
    gets 5 more entries."""
    A = np.zeros((m, n))
    for _ in range(k):
        q = min(m, n)
        rows, cols = rng.permutation(m)[:q], rng.permutation(n)[:q]
        keep = rng.random(q) < 0.5
        A[rows[keep], cols[keep]] = rng.standard_normal(int(keep.sum()))
    if wide_row:
        A[m // 2, rng.choice(n, size=5, replace=False)] = rng.standard_normal(5)
    return sp.csr_matrix(A)


@pytest.mark.parametrize("case", ["sparse256", "dense_rect", "mixed", "long_rows1024", "short_rows",
                                  "short_rows_two_per_thread", "short_rows_but_one"])
def test_norm2_both_forms_bitexact(gpu, power_form, case):
    # short_rows*: every operator of the F^T F chain has rows of at most 4
    # entries (the persistent kernel's branch-free round loop), on vectors of
    # up to 1024 and (two rows per thread) 1500 entries; short_rows_but_one:
    # one row of 5 entries sends the same kind of chain through the general loop
    rng = np.random.default_rng(17)
    if case.startswith("short_rows"):
        wide = case.endswith("but_one")
        if case == "short_rows_two_per_thread":
            facs = [_short_factor(1500, 1500, rng, 4) for _ in range(3)]
        else:
            facs = [_short_factor(700, 600, rng, 4), _short_factor(600, 650, rng, 2, wide_row=wide),
                    _short_factor(650, 700, rng, 4)]
        for f in facs:
            rl = np.diff(f.indptr)
            cl = np.diff(sp.csc_matrix(f).indptr)
            assert (max(rl.max(), cl.max()) <= 4) or wide
    elif case == "sparse256":
        facs = _csr_chain_factors(n=256, nnz=4, J=4)
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


def _launches(F):
    import ctypes as C
    from faust_amd import _lib
    n = C.c_int()
    _lib.check(_lib.lib().fh_faust_last_launches(F.h, C.byref(n)))
    return n.value


@pytest.mark.parametrize("wc", ["auto", "1", "2", "4", "8"])
@pytest.mark.parametrize("op", ["N", "T"])
def test_fused_chain_matches_per_factor(gpu, monkeypatch, wc, op):
    # the one-launch FP64 chain (column panels in LDS) against the
    # factor-by-factor launches and the oracle, bit for bit: rectangular
    # factors, an empty row, column counts that leave partial panels
    if wc != "auto":
        monkeypatch.setenv("FH_CHAIN_WC", wc)
    rng = np.random.default_rng(21)
    shapes = [(300, 700), (700, 129), (129, 1000), (1000, 500)]
    facs = []
    for r, c in shapes:
        a = sp.random(r, c, density=min(1.0, 9.0 / c), random_state=rng, format="csr")
        a = sp.csr_matrix(a.toarray())
        facs.append(a)
    facs[1] = sp.csr_matrix(np.where(np.arange(700)[:, None] == 4, 0.0, facs[1].toarray()))  # empty row 4
    F = faust_amd.Faust(facs)
    G = F if op == "N" else F.T
    rows_in = shapes[-1][1] if op == "N" else shapes[0][0]
    for p in (1, 7, 64, 200):
        X = np.asfortranarray(rng.standard_normal((rows_in, p)))
        monkeypatch.setenv("FH_CHAIN_FUSED", "1")
        Y1 = G @ X
        assert _launches(G) == 1
        monkeypatch.setenv("FH_CHAIN_FUSED", "0")
        Y2 = G @ X
        assert _launches(G) == len(facs)
        assert np.array_equal(Y1, Y2)
        assert np.array_equal(Y1, oracle.chain_multiply(facs, X, op))


@pytest.mark.parametrize("wc", ["1", "2", "4"])
@pytest.mark.parametrize("op", ["N", "T"])
def test_prefetched_chain_matches_per_factor(gpu, monkeypatch, wc, op):
    # the one-launch chain with one thread per row (every row <= 1024) and the
    # entries loaded two factors ahead, against the factor-by-factor launches
    # and the oracle, bit for bit: rectangular factors, an empty row, rows of
    # 0 to 40 entries (past the 8 / 16 prefetched: the loop reads the rest),
    # odd chain lengths, column counts that leave partial panels
    monkeypatch.setenv("FH_CHAIN_PF_WC", wc)
    monkeypatch.setenv("FH_AUTOTUNE", "0")  # the default form, not the measured one
    monkeypatch.delenv("FH_CHAIN_FUSED", raising=False)
    rng = np.random.default_rng(23)
    for shapes in ([(300, 700), (700, 129), (129, 1000), (1000, 500)],
                   [(1024, 1024)] * 5,
                   [(64, 1024), (1024, 1024), (1024, 17)]):
        facs = []
        for r, c in shapes:
            a = sp.random(r, c, density=min(1.0, 9.0 / c), random_state=rng, format="csr").toarray()
            a[3, : min(c, 40)] = rng.standard_normal(min(c, 40))  # a row of 40 entries
            a[min(4, r - 1), :] = 0.0  # an empty row
            facs.append(sp.csr_matrix(a))
        F = faust_amd.Faust(facs)
        G = F if op == "N" else F.T
        rows_in = shapes[-1][1] if op == "N" else shapes[0][0]
        for p in (1, 7, 64, 200):
            X = np.asfortranarray(rng.standard_normal((rows_in, p)))
            monkeypatch.delenv("FH_CHAIN_PF", raising=False)
            Y1 = G @ X
            assert _launches(G) == 1
            monkeypatch.setenv("FH_CHAIN_PF", "0")
            monkeypatch.setenv("FH_CHAIN_FUSED", "0")
            Y2 = G @ X
            monkeypatch.delenv("FH_CHAIN_FUSED")
            assert _launches(G) == len(facs)
            assert np.array_equal(Y1, Y2)
            assert np.array_equal(Y1, oracle.chain_multiply(facs, X, op))


@pytest.mark.parametrize("op", ["N", "T"])
def test_autotuned_chain_form(gpu, monkeypatch, op):
    # AUTO mode measures the forms of an FP64 CSR chain on its first product
    # and keeps the quickest: every product, before and after, has the
    # oracle's bits; clearing the measurements measures again; FH_AUTOTUNE=0
    # gives the default (one-launch) form
    from faust_amd import _lib
    monkeypatch.delenv("FH_CHAIN_FUSED", raising=False)
    monkeypatch.delenv("FH_CHAIN_PF", raising=False)
    monkeypatch.delenv("FH_AUTOTUNE", raising=False)
    facs = _csr_chain_factors()
    F = faust_amd.Faust(facs)
    G = F if op == "N" else F.T
    X = np.asfortranarray(np.random.default_rng(9).standard_normal((1024, 48)))
    ref = oracle.chain_multiply(facs, X, op)
    _lib.check(_lib.lib().fh_autotune_clear())
    for _ in range(3):
        assert np.array_equal(G @ X, ref)
        assert _launches(G) in (1, len(facs))
    _lib.check(_lib.lib().fh_autotune_clear())
    assert np.array_equal(G @ X, ref)
    monkeypatch.setenv("FH_AUTOTUNE", "0")
    assert np.array_equal(G @ X, ref)
    assert _launches(G) == 1


def test_fused_chain_falls_back_when_too_tall(gpu):
    # 12000-row intermediates do not fit two LDS buffers even at one column
    rng = np.random.default_rng(22)
    facs = [faust_amd.rand_csr(12000, 12000, 3, np.random.default_rng(22 + f)) for f in range(3)]
    F = faust_amd.Faust(facs)
    X = np.asfortranarray(rng.standard_normal((12000, 4)))
    Y = F @ X
    assert _launches(F) == 3
    assert np.array_equal(Y, oracle.chain_multiply(facs, X, "N"))


def test_long_chains_fuse_without_ell_views(gpu, monkeypatch):
    # the 10-factor Hadamard Faust of order 1024 (2 nnz per row and column):
    # one launch (entries loaded ahead) by default; without that form, one
    # launch per factor through the ELL views, one launch for the whole chain
    # (staged factors) when the views are off; the 5-factor chain of 1024^2
    # factors is then never fused.  Same bits every way.
    monkeypatch.delenv("FH_CHAIN_FUSED", raising=False)
    monkeypatch.delenv("FH_CHAIN_WC", raising=False)
    monkeypatch.setenv("FH_AUTOTUNE", "0")
    facs = [sp.csr_matrix(f) for f in faust_amd.wht_factors(1024)]
    F = faust_amd.Faust(facs)
    X = np.asfortranarray(np.random.default_rng(7).standard_normal((1024, 33)))
    for pf, ell, launches in (("1", "1", 1), ("0", "1", 10), ("0", "0", 1)):
        monkeypatch.setenv("FH_CHAIN_PF", pf)
        monkeypatch.setenv("FH_CSR_ELL", ell)
        for G, op in ((F, "N"), (F.T, "T")):
            Y = G @ X
            assert _launches(G) == launches
            assert np.array_equal(Y, oracle.chain_multiply(facs, X, op))
    F5 = faust_amd.Faust(_csr_chain_factors())
    for pf, launches in (("1", 1), ("0", 5)):
        monkeypatch.setenv("FH_CHAIN_PF", pf)
        F5 @ X
        assert _launches(F5) == launches


def test_optimize_time_fp64(gpu):
    # Faust.optimize_time: time the execution forms, keep the quickest; every
    # form of an FP64 chain gives the same bits
    import ctypes as C
    from faust_amd import _lib
    facs = _csr_chain_factors()
    F = faust_amd.Faust(facs)
    X = np.asfortranarray(np.random.default_rng(8).standard_normal((1024, 64)))
    ref = oracle.chain_multiply(facs, X, "N")
    G = F.optimize_time(nsamples=3, mat=X)
    assert G is not F and G.mul_mode_name in ("AUTO", "PER_FACTOR", "FUSED")
    assert set(G.mul_mode_times_ms) == {"AUTO", "PER_FACTOR", "FUSED"}
    assert np.array_equal(G @ X, ref)
    m = C.c_int()
    _lib.check(_lib.lib().fh_faust_get_mul_mode(F.h, C.byref(m)))
    assert m.value == _lib.FH_MUL_AUTO  # F itself unchanged
    for mode, n in ((_lib.FH_MUL_FUSED, 1), (_lib.FH_MUL_PER_FACTOR, 5)):
        F._set_mul_mode(mode)
        assert np.array_equal(F @ X, ref)
        assert _launches(F) == n
        Ft = F.T  # the mode follows .T
        assert np.array_equal(Ft @ X, oracle.chain_multiply(facs, X, "T"))
        assert _launches(Ft) == n
    H = F.optimize_time(transp=True, inplace=True)
    assert H is F
    with pytest.raises(faust_amd.FaustHipError):
        F._set_mul_mode(7)


def test_get_product_fp32_chain(gpu):
    """The FP64 get_product refuses FP32 chains; Faust.toarray forms their
    product as F times the identity, in FP32."""
    import ctypes as C
    import torch
    from faust_amd import _lib
    rng = np.random.default_rng(12)
    facs = [sp.csr_matrix(rng.standard_normal((48, 48)) * (rng.random((48, 48)) < 0.2)).astype(np.float32)
            for _ in range(3)]
    F = faust_amd.Faust(facs, dtype=np.float32)
    out = torch.empty((48, 48), dtype=torch.float64, device="cuda")
    rc = _lib.lib().fh_faust_get_product(F.h, C.c_void_p(out.data_ptr()), None)
    assert rc == 4  # FH_ERR_UNSUPPORTED
    ref = facs[0].toarray().astype(np.float64) @ facs[1].toarray() @ facs[2].toarray()
    assert np.linalg.norm(F.toarray() - ref) <= 1e-5 * np.linalg.norm(ref)


def test_spectral_norm_fp32_chain(gpu):
    """The FP64 power iteration refuses FP32 chains; Faust.norm(2) runs it on
    an FP64 copy of their factors."""
    import ctypes as C
    from faust_amd import _lib
    rng = np.random.default_rng(13)
    facs = [sp.csr_matrix(rng.standard_normal((40, 40)) * (rng.random((40, 40)) < 0.3)).astype(np.float32)
            for _ in range(3)]
    F = faust_amd.Faust(facs, dtype=np.float32)
    nrm, flag = C.c_double(), C.c_int()
    assert _lib.lib().fh_faust_spectral_norm(F.h, 100, 1e-6, C.byref(nrm), C.byref(flag), None) == 4
    ref = np.linalg.norm(facs[0].toarray().astype(np.float64) @ facs[1].toarray() @ facs[2].toarray(), 2)
    assert abs(F.norm(2) - ref) <= 1e-4 * ref
    # a dense FP32 factor too (no nonzero count is taken of it)
    G = faust_amd.Faust([facs[0].toarray(), facs[1]], dtype=np.float32)
    ref2 = np.linalg.norm(facs[0].toarray().astype(np.float64) @ facs[1].toarray(), 2)
    assert abs(G.norm(2) - ref2) <= 1e-4 * ref2


def test_nnz_of_dense_factors_all_dtypes(gpu):
    """fh_mat_info counts the nonzeros of dense FP64, FP32 and BF16 factors."""
    rng = np.random.default_rng(14)
    a = rng.standard_normal((37, 29)) * (rng.random((37, 29)) < 0.4)
    want = int(np.count_nonzero(a))
    for dt in (np.float64, np.float32, "bf16"):
        f = faust_amd.Factor(a, dt)
        assert f.info()[4] == want, dt


@pytest.mark.parametrize("dense", [False, True])
def test_spectral_norm_beyond_2048_bitexact(gpu, dense):
    """Operators wider than 2048 go through the wide power iteration with x
    read from global memory; same fma chains, so the oracle's bits."""
    rng = np.random.default_rng(15)
    if dense:
        n = 2304
        facs = [_csr_chain_factors(seed=0xD3, n=n, nnz=6, J=1)[0], np.asfortranarray(rng.standard_normal((n, n)) / 48.0),
                _csr_chain_factors(seed=0xD4, n=n, nnz=6, J=1)[0]]
    else:
        facs = _csr_chain_factors(seed=0xB16, n=4096, nnz=8, J=3)
    F = faust_amd.Faust(facs)
    nrm, _ = oracle.spectral_norm(facs)
    assert F.norm(2) == nrm


@pytest.mark.parametrize("widest", [12, 30])
@pytest.mark.parametrize("op", ["N", "T"])
def test_fp64_ell_view_same_bits(gpu, monkeypatch, op, widest):
    """FP64 CSR factors with rows of at most 32 entries go through their ELL
    view (8 to 32 slots): the oracle's bits, with and without it (ragged and
    empty rows)."""
    rng = np.random.default_rng(61 + widest)
    facs = []
    for s in [(200, 150), (150, 170), (170, 200)]:
        a = np.zeros(s)
        for r in range(s[0]):
            k = rng.integers(0, widest + 1)
            a[r, rng.choice(s[1], size=k, replace=False)] = rng.standard_normal(k)
        facs.append(sp.csr_matrix(a))
    F = faust_amd.Faust(facs)
    F._set_mul_mode(1)  # one launch per factor (the fused chain reads CSR)
    rows_in = facs[-1].shape[1] if op == "N" else facs[0].shape[0]
    X = np.asfortranarray(rng.standard_normal((rows_in, 37)))
    ref = oracle.chain_multiply(facs, X, op)
    for ell in ("1", "0"):
        monkeypatch.setenv("FH_CSR_ELL", ell)
        Y = (F @ X) if op == "N" else (F.T @ X)
        assert np.array_equal(Y, ref), ell


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_edge_shapes(gpu, dtype):
    """A factor with no nonzeros, a single-factor Faust, a 1-row factor and
    X with no columns, in FP64 and FP32 chains, N and T."""
    rng = np.random.default_rng(70)
    tol = 0.0 if dtype == np.float64 else 1e-5

    def check(facs):
        F = faust_amd.Faust(facs, dtype=dtype)
        for op in ("N", "T"):
            G = F if op == "N" else F.T
            rows_in = facs[-1].shape[1] if op == "N" else facs[0].shape[0]
            for p in (0, 1, 7):
                X = rng.standard_normal((rows_in, p))
                if dtype == np.float32:
                    X = X.astype(np.float32).astype(np.float64)
                Y = G @ X
                ref = oracle.chain_multiply(facs, np.asfortranarray(X), op) if p else np.zeros((G.shape[0], 0))
                assert Y.shape == ref.shape
                if tol == 0.0:
                    assert np.array_equal(Y, ref)
                else:
                    assert np.linalg.norm(Y - ref) <= tol * max(np.linalg.norm(ref), 1e-30)

    a = sp.csr_matrix(rng.standard_normal((30, 20)) * (rng.random((30, 20)) < 0.2)).astype(dtype).astype(np.float64)
    b = sp.csr_matrix(rng.standard_normal((20, 25)) * (rng.random((20, 25)) < 0.2)).astype(dtype).astype(np.float64)
    zero = sp.csr_matrix((20, 20))
    one_row = sp.csr_matrix(rng.standard_normal((1, 30)) * (rng.random((1, 30)) < 0.5)).astype(dtype).astype(np.float64)
    check([a, zero, b])      # an all-zero factor: the product is zero
    check([a])               # one factor
    check([one_row, a, b])   # a 1 x 30 first factor
