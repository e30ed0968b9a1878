"""FP32 / BF16 Faust x dense (BSR on MFMA, FP32 CSR) vs the FP64 oracle.

Tolerances: BF16 activations are rounded to nearest-even after every factor
(relative 2^-9 per entry); a single BSR product has max |err| <= 2^-8 * |y|
+ 1e-3 * ||y||_inf (FP32 accumulation over 128 terms, then one BF16
rounding).  FP32 block values enter as a BF16 hi/lo pair (2^-16 relative).
FP32 CSR chains are compared with relative Frobenius error <= 1e-5."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
import faust_amd
from faust_amd.bsr import bf16_bits_to_f32, f32_to_bf16_bits, rand_bsr

pytestmark = pytest.mark.gpu


def bf16_round(a):
    return bf16_bits_to_f32(f32_to_bf16_bits(np.asarray(a, dtype=np.float32))).astype(np.float64)


def _bsr_rounded(B):
    B = sp.bsr_matrix(B)
    return sp.bsr_matrix((bf16_round(B.data), B.indices, B.indptr), shape=B.shape)


@pytest.mark.parametrize("op", ["N", "T"])
@pytest.mark.parametrize("p", [512, 96, 8])
def test_bsr_bf16_single_product(gpu, op, p):
    rng = np.random.default_rng(1)
    B = _bsr_rounded(rand_bsr(1024, 768, 32, 4, rng, scale=1 / np.sqrt(128)))
    rows_in = B.shape[1] if op == "N" else B.shape[0]
    X = bf16_round(rng.standard_normal((rows_in, p)))
    F = faust_amd.Faust([B], dtype="bf16")
    Y = (F @ X) if op == "N" else (F.T @ X)
    ref = oracle.chain_multiply([B], X, op)
    tol = 2.0 ** -8 * np.abs(ref) + 1e-3 * np.abs(ref).max()
    assert np.all(np.abs(Y - ref) <= tol)


def test_bsr_fp32_split_single_product(gpu):
    rng = np.random.default_rng(2)
    B = rand_bsr(512, 512, 32, 4, rng, scale=1 / np.sqrt(128))
    B32 = sp.bsr_matrix((B.data.astype(np.float32).astype(np.float64), B.indices, B.indptr), shape=B.shape)
    X = bf16_round(rng.standard_normal((512, 64)))
    Y = faust_amd.Faust([B32], dtype=np.float32) @ X
    ref = oracle.chain_multiply([B32], X, "N")
    tol = 2.0 ** -8 * np.abs(ref) + 1e-4 * np.abs(ref).max()
    assert np.all(np.abs(Y - ref) <= tol)
    # the split keeps FP32 accuracy of the block values: finer than a BF16 factor
    Yb = faust_amd.Faust([B32], dtype="bf16") @ X
    assert np.linalg.norm(Y - ref) < np.linalg.norm(Yb - ref)


def test_bsr_chain_32x32_blocks_error(gpu):
    # 3 factors 2048^2, 32x32 blocks, 4 blocks per block row; error vs FP64 oracle (same BF16 inputs) stated
    rng = np.random.default_rng(3)
    facs = [_bsr_rounded(rand_bsr(2048, 2048, 32, 4, rng, scale=1 / np.sqrt(128))) for _ in range(3)]
    X = bf16_round(rng.standard_normal((2048, 512)))
    Y = faust_amd.Faust(facs, dtype="bf16") @ X
    ref = oracle.chain_multiply(facs, X, "N")
    rel = np.linalg.norm(Y - ref) / np.linalg.norm(ref)
    assert rel < 8e-3, rel


def test_bsr_other_block_size_goes_csr(gpu):
    rng = np.random.default_rng(4)
    B = rand_bsr(96, 64, 16, 2, rng)
    B32 = sp.bsr_matrix((B.data.astype(np.float32).astype(np.float64), B.indices, B.indptr), shape=B.shape)
    X = rng.standard_normal((64, 12)).astype(np.float32).astype(np.float64)
    Y = faust_amd.Faust([B32], dtype=np.float32) @ X
    ref = oracle.chain_multiply([B32], X, "N")
    assert np.linalg.norm(Y - ref) / np.linalg.norm(ref) < 1e-6


@pytest.mark.parametrize("op", ["N", "T"])
@pytest.mark.parametrize("p", [256, 13, 1])
def test_csr_fp32_chain(gpu, op, p):
    rng = np.random.default_rng(5)
    facs = [faust_amd.rand_csr(2048, 2048, 12, rng).astype(np.float32).astype(np.float64) for _ in range(4)]
    X = rng.standard_normal((2048, p)).astype(np.float32).astype(np.float64)
    F = faust_amd.Faust(facs, dtype=np.float32)
    Y = (F @ X) if op == "N" else (F.T @ X)
    ref = oracle.chain_multiply(facs, X, op)
    assert np.linalg.norm(Y - ref) / np.linalg.norm(ref) < 1e-5


def test_csr_fp32_rectangular_chain(gpu):
    rng = np.random.default_rng(6)
    shapes = [(100, 64), (64, 200), (200, 36)]
    facs = [sp.csr_matrix(rng.standard_normal(s) * (rng.random(s) < 0.2)).astype(np.float32).astype(np.float64)
            for s in shapes]
    X = rng.standard_normal((36, 7)).astype(np.float32).astype(np.float64)
    F = faust_amd.Faust(facs, dtype=np.float32)
    ref = oracle.chain_multiply(facs, X, "N")
    assert np.linalg.norm(F @ X - ref) / np.linalg.norm(ref) < 1e-5
    X2 = rng.standard_normal((100, 5)).astype(np.float32).astype(np.float64)
    ref2 = oracle.chain_multiply(facs, X2, "T")
    assert np.linalg.norm(F.T @ X2 - ref2) / np.linalg.norm(ref2) < 1e-5


def test_bsr_to_dense(gpu):
    rng = np.random.default_rng(7)
    B = _bsr_rounded(rand_bsr(128, 96, 32, 2, rng))
    f = faust_amd.Factor(B, "bf16")
    assert np.array_equal(f.toarray(), B.toarray())


def _ragged_bsr(rng, nbr, nbc, bs=32, wide_row=None):
    # block rows with 0..6 blocks (some empty); optionally one block row with
    # `wide_row` blocks, wider than the ELL slot limit
    rows = []
    for br in range(nbr):
        k = int(rng.integers(0, 7)) if br != 3 else 0
        if wide_row and br == 5:
            k = wide_row
        rows.append(np.sort(rng.choice(nbc, size=min(k, nbc), replace=False)))
    indptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])])
    idx = np.concatenate(rows).astype(np.int64)
    data = rng.standard_normal((len(idx), bs, bs)) / np.sqrt(128)
    return sp.bsr_matrix((data, idx, indptr), shape=(nbr * bs, nbc * bs))


@pytest.mark.parametrize("kernel", ["0", "1", "2", "3", "4", "5"])
@pytest.mark.parametrize("wide", [None, 20])
@pytest.mark.parametrize("op", ["N", "T"])
@pytest.mark.parametrize("xsplit_nt", ["1-0", "2-1"])
def test_bsr_kernel_variants(gpu, monkeypatch, kernel, wide, op, xsplit_nt):
    # row-pointer kernel, ELL with 32 / 64 / 128 columns per wave (with the XCD
    # split of the block rows and streaming loads/stores), the K-split ELL
    # kernel (two block rows per workgroup: the odd block-row counts 25 leave
    # one row alone) and the shared-block kernel (4 column tiles per
    # workgroup: p = 200 leaves a partial group; rows of up to 6 blocks make
    # two groups of 4 slots); ragged block rows, empty block rows and (wide) a block
    # row beyond the ELL limit; a 3-factor chain exercises the tiled
    # intermediates
    monkeypatch.setenv("FH_BSR_KERNEL", kernel)
    monkeypatch.setenv("FH_BSR_XSPLIT", xsplit_nt[0])
    monkeypatch.setenv("FH_BSR_NT", xsplit_nt[2])
    rng = np.random.default_rng(11)
    facs = [_bsr_rounded(_ragged_bsr(rng, r, c, wide_row=wide)) for r, c in ((25, 24), (24, 25), (25, 24))]
    for p in (200, 8):
        X = bf16_round(rng.standard_normal((32 * (24 if op == "N" else 25), p)))
        F = faust_amd.Faust(facs, dtype="bf16")
        Y = (F @ X) if op == "N" else (F.T @ X)
        ref = oracle.chain_multiply(facs, X, op)
        rel = np.linalg.norm(Y - ref) / np.linalg.norm(ref)
        assert rel < 8e-3, rel
        # the tiled intermediates hold the same BF16 values as column-major
        # ones: the chain equals the factors applied one at a time, bit for bit
        Z = X
        for f in (facs[::-1] if op == "N" else facs):
            G = faust_amd.Faust([f], dtype="bf16")
            Z = (G @ Z) if op == "N" else (G.T @ Z)
        assert np.array_equal(Y, Z)
        Y1 = faust_amd.Faust([facs[-1]], dtype="bf16") @ X if op == "N" else faust_amd.Faust([facs[0]], dtype="bf16").T @ X
        r1 = oracle.chain_multiply([facs[-1] if op == "N" else facs[0]], X, op)
        assert np.all(np.abs(Y1 - r1) <= 2.0 ** -8 * np.abs(r1) + 1e-3 * np.abs(r1).max())


@pytest.mark.parametrize("rpt", ["1", "2"])
@pytest.mark.parametrize("op", ["N", "T"])
def test_csr_fp32_rows_per_thread(gpu, monkeypatch, rpt, op):
    monkeypatch.setenv("FH_CSR_RPT", rpt)
    rng = np.random.default_rng(12)
    facs = []
    for s in [(300, 250), (250, 333), (333, 300)]:
        a = rng.standard_normal(s) * (rng.random(s) < rng.uniform(0.0, 0.15, size=(s[0], 1)))
        a[7] = 0.0  # an empty row
        facs.append(sp.csr_matrix(a).astype(np.float32).astype(np.float64))
    for p in (70, 33, 1):
        rows_in = facs[-1].shape[1] if op == "N" else facs[0].shape[0]
        X = rng.standard_normal((rows_in, p)).astype(np.float32).astype(np.float64)
        F = faust_amd.Faust(facs, dtype=np.float32)
        Y = (F @ X) if op == "N" else (F.T @ X)
        ref = oracle.chain_multiply(facs, X, op)
        assert np.linalg.norm(Y - ref) / np.linalg.norm(ref) < 1e-5


@pytest.mark.parametrize("dtype", ["bf16", np.float32])
def test_bsr_chain_launch_count(gpu, dtype):
    # one product kernel per factor
    import ctypes as C
    from faust_amd import _lib
    rng = np.random.default_rng(13)
    facs = [rand_bsr(512, 512, 32, 4, rng, scale=1 / np.sqrt(128)) for _ in range(3)]
    F = faust_amd.Faust(facs, dtype=dtype)
    F @ bf16_round(rng.standard_normal((512, 64)))
    n = C.c_int()
    _lib.check(_lib.lib().fh_faust_last_launches(F.h, C.byref(n)))
    assert n.value == 3


@pytest.mark.parametrize("dtype", ["bf16", np.float32])
@pytest.mark.parametrize("op", ["N", "T"])
def test_bsr_persistent_chain_matches_per_factor(gpu, monkeypatch, dtype, op):
    import ctypes as C
    from faust_amd import _lib
    rng = np.random.default_rng(13)
    facs = [rand_bsr(2048, 2048, 32, 4, rng, scale=1 / np.sqrt(128)) for _ in range(3)]
    if dtype == "bf16":
        facs = [_bsr_rounded(f) for f in facs]
    F = faust_amd.Faust(facs, dtype=dtype)
    G = F if op == "N" else F.T
    for p in (512, 200, 8):
        X = bf16_round(rng.standard_normal((2048, p)))
        # the one-launch chain (k_bsr_chain.hip), every call checked
        monkeypatch.setenv("FH_BSR_CHAIN", "1")
        monkeypatch.setenv("FH_BSR_CHAIN_CHECK", "1")
        Y1 = G @ X
        n = C.c_int()
        _lib.check(_lib.lib().fh_faust_last_launches(G.h, C.byref(n)))
        assert n.value == 1
        monkeypatch.setenv("FH_BSR_CHAIN", "0")
        monkeypatch.setenv("FH_BSR_KERNEL", "2")
        Y2 = G @ X
        monkeypatch.delenv("FH_BSR_KERNEL")
        assert np.array_equal(Y1, Y2)


def test_bsr_8192x6_full_size_against_oracle_columns(gpu):
    # the benchmarked Faust itself (6 BSR factors 8192^2, 4 blocks per block
    # row, BF16): the full product on X 8192x512, its first 16 columns against
    # the FP64 oracle, and linearity of the whole product
    rng = np.random.default_rng(0xB5B)
    facs = [_bsr_rounded(rand_bsr(8192, 8192, 32, 4, np.random.default_rng(0xB5B + f), scale=1 / np.sqrt(128)))
            for f in range(6)]
    X = bf16_round(rng.standard_normal((8192, 512)))
    F = faust_amd.Faust(facs, dtype="bf16")
    Y = F @ X
    ref = oracle.chain_multiply(facs, X[:, :16], "N")
    rel = np.linalg.norm(Y[:, :16] - ref) / np.linalg.norm(ref)
    assert rel < 1.2e-2, rel  # six BF16 roundings of the activations
    # columns are independent: the same columns computed alone give the same bits
    Y8 = F @ np.asfortranarray(X[:, 100:108])
    assert np.array_equal(Y8, Y[:, 100:108])


def test_csr_fp32_16384x6_full_size_against_oracle_columns(gpu):
    rng = np.random.default_rng(0xC5F)
    facs = [faust_amd.rand_csr(16384, 16384, 12, np.random.default_rng(0xC5F + f)).astype(np.float32)
            .astype(np.float64) for f in range(6)]
    X = rng.standard_normal((16384, 256)).astype(np.float32).astype(np.float64)
    F = faust_amd.Faust(facs, dtype=np.float32)
    Y = F @ X
    ref = oracle.chain_multiply(facs, X[:, :8], "N")
    assert np.linalg.norm(Y[:, :8] - ref) / np.linalg.norm(ref) < 1e-5
    Yt = F.T @ X[:, :8]
    reft = oracle.chain_multiply(facs, X[:, :8], "T")
    assert np.linalg.norm(Yt - reft) / np.linalg.norm(reft) < 1e-5


def test_optimize_time_bsr(gpu):
    rng = np.random.default_rng(14)
    facs = [_bsr_rounded(rand_bsr(1024, 1024, 32, 4, np.random.default_rng(140 + f), scale=1 / np.sqrt(128)))
            for f in range(3)]
    F = faust_amd.Faust(facs, dtype="bf16")
    X = bf16_round(rng.standard_normal((1024, 256)))
    G = F.optimize_time(nsamples=2, mat=X)
    # five per-factor kernels and the one-launch chain
    assert len(G.mul_mode_times_ms) == 6 and "FUSED" in G.mul_mode_times_ms
    assert G.mul_mode_name.startswith("BSR_KERNEL_") or G.mul_mode_name == "FUSED"
    ref = oracle.chain_multiply(facs, X, "N")
    for H in (G, F):
        Y = H @ X
        assert np.linalg.norm(Y - ref) / np.linalg.norm(ref) < 8e-3


@pytest.mark.parametrize("width", [3, 8, 11, 16])
@pytest.mark.parametrize("op", ["N", "T"])
def test_csr_fp32_ell_view_same_bits(gpu, monkeypatch, width, op):
    """The ELL view of FP32 CSR factors (rows of at most 16 entries) gives the
    CSR kernel's bits: the same terms in the same order, unused slots not
    accumulated.  Ragged rows, empty rows, column-major and row-major outputs."""
    rng = np.random.default_rng(40 + width)
    facs = []
    for s in [(300, 256), (256, 333), (333, 300)]:
        a = np.zeros(s)
        for r in range(s[0]):
            k = rng.integers(0, width + 1)  # 0..width entries, some rows empty
            a[r, rng.choice(s[1], size=k, replace=False)] = rng.standard_normal(k)
        facs.append(sp.csr_matrix(a).astype(np.float32).astype(np.float64))
    if op == "T":  # the transpose's rows are the columns: keep them short too
        facs = [sp.csr_matrix(f.T.toarray()) for f in reversed(facs)]
        facs = [sp.csr_matrix(f.T.toarray()) for f in reversed(facs)]
    F = faust_amd.Faust(facs, dtype=np.float32)
    for p in (70, 33, 1):
        rows_in = facs[-1].shape[1] if op == "N" else facs[0].shape[0]
        X = rng.standard_normal((rows_in, p)).astype(np.float32).astype(np.float64)
        monkeypatch.setenv("FH_CSR_ELL", "1")
        Y1 = (F @ X) if op == "N" else (F.T @ X)
        monkeypatch.setenv("FH_CSR_ELL", "0")
        Y0 = (F @ X) if op == "N" else (F.T @ X)
        assert np.array_equal(Y1, Y0)
        ref = oracle.chain_multiply(facs, X, op)
        assert np.linalg.norm(Y1 - ref) / np.linalg.norm(ref) < 1e-5


@pytest.mark.parametrize("op", ["N", "T"])
def test_fp32_chain_with_dense_factors(gpu, op):
    """Dense FP32 factors in an FP32 chain run as CSR twins of themselves."""
    rng = np.random.default_rng(50)
    facs = [sp.csr_matrix(rng.standard_normal((96, 80)) * (rng.random((96, 80)) < 0.1)).astype(np.float32)
            .astype(np.float64), np.asfortranarray(rng.standard_normal((80, 64)).astype(np.float32).astype(np.float64)),
            sp.csr_matrix(rng.standard_normal((64, 72)) * (rng.random((64, 72)) < 0.1)).astype(np.float32)
            .astype(np.float64)]
    F = faust_amd.Faust(facs, dtype=np.float32)
    for p in (33, 1):
        rows_in = facs[-1].shape[1] if op == "N" else facs[0].shape[0]
        X = rng.standard_normal((rows_in, p)).astype(np.float32).astype(np.float64)
        Y = (F @ X) if op == "N" else (F.T @ X)
        ref = oracle.chain_multiply(facs, X, op)
        assert np.linalg.norm(Y - ref) / np.linalg.norm(ref) < 1e-5


@pytest.mark.parametrize("dtype", ["bf16", np.float32])
def test_bsr_chain_with_empty_factor(gpu, dtype):
    """A BSR factor with no blocks (every ELL slot a zero block) in a chain:
    the product is zero; with it removed the chain matches the oracle."""
    rng = np.random.default_rng(71)
    A = rand_bsr(256, 256, 32, 3, rng)
    Z = sp.bsr_matrix((np.zeros((0, 32, 32)), np.zeros(0, dtype=np.int32), np.zeros(9, dtype=np.int32)),
                      shape=(256, 256))
    X = rng.standard_normal((256, 64))
    F = faust_amd.Faust([A, Z, A], dtype=dtype)
    assert np.all(F @ X == 0.0)
    G = faust_amd.Faust([A, A], dtype=dtype)
    ref = oracle.chain_multiply([sp.csr_matrix(bf16_bits_to_f32(f32_to_bf16_bits(A.toarray())).astype(np.float64))
                                 if dtype == "bf16" else sp.csr_matrix(A)] * 2, np.asfortranarray(X), "N")
    assert np.linalg.norm(G @ X - ref) <= 2e-2 * np.linalg.norm(ref)  # BF16 activations


def _f32(B):
    B = sp.bsr_matrix(B)
    return sp.bsr_matrix((B.data.astype(np.float32).astype(np.float64), B.indices, B.indptr), shape=B.shape)


@pytest.mark.parametrize("op", ["N", "T"])
@pytest.mark.parametrize("p", [512, 40, 1])
def test_bsr_fp32_activations_single_product(gpu, op, p):
    # FP32 blocks x FP32 X on v_mfma_f32_32x32x2_f32 (MatBSR<float>::multiply's
    # precision): FP32 rounding of products summed over 96 terms, each
    # |err| <= 96 * 2^-24 * sum |b_ik x_kj| (a standard FP32 dot-product bound)
    rng = np.random.default_rng(21)
    B = _f32(rand_bsr(1024, 768, 32, 3, rng, scale=1 / np.sqrt(96)))
    rows_in = B.shape[1] if op == "N" else B.shape[0]
    X = rng.standard_normal((rows_in, p)).astype(np.float32).astype(np.float64)
    F = faust_amd.Faust([B], dtype=np.float32, act_dtype=np.float32)
    assert F.act_dtype == torch.float32
    Y = (F @ X) if op == "N" else (F.T @ X)
    ref = oracle.chain_multiply([B], X, op)
    absref = oracle.chain_multiply([abs(B)], np.abs(X), op)
    assert np.all(np.abs(Y - ref) <= 100 * 2.0 ** -24 * absref + 1e-30)


@pytest.mark.parametrize("bpr", [3, 4, 5, 9])
@pytest.mark.parametrize("op", ["N", "T"])
def test_bsr_fp32_shared_block_kernel_same_bits(gpu, monkeypatch, bpr, op):
    # k_bsr_f32s (a row's blocks staged in LDS for four column tiles, rounds of
    # 4 blocks, the rest one by one) against k_bsr_f32 (FH_BSR_F32_KERNEL=0):
    # the same MFMA sequence per output tile, the same bits; p = 200 leaves a
    # partial 128-column group
    rng = np.random.default_rng(40 + bpr)
    B = _f32(rand_bsr(512, 384, 32, bpr, rng, scale=1 / np.sqrt(32 * bpr)))
    rows_in = B.shape[1] if op == "N" else B.shape[0]
    X = rng.standard_normal((rows_in, 200)).astype(np.float32).astype(np.float64)
    F = faust_amd.Faust([B], dtype=np.float32, act_dtype=np.float32)
    G = F if op == "N" else F.T
    monkeypatch.delenv("FH_BSR_F32_KERNEL", raising=False)
    Y1 = G @ X
    monkeypatch.setenv("FH_BSR_F32_KERNEL", "0")
    Y0 = G @ X
    assert np.array_equal(Y1, Y0)
    ref = oracle.chain_multiply([B], X, op)
    assert np.linalg.norm(Y1 - ref) / np.linalg.norm(ref) < 1e-6


def test_bsr_fp32_chain_8192x6_full_size_against_oracle_columns(gpu):
    # config 3a's FP32 leg at full size with FP32 activations: 6 FP32 BSR
    # factors 8192^2 (32x32 blocks, 4 per block row, N(0, 1/128)), X 8192x512
    # FP32; 16 columns against the FP64 oracle.  FP32 level: relative
    # Frobenius error below 1e-5 (the BF16-activation form is at ~1e-2)
    rng = np.random.default_rng(0xF32)
    facs = [_f32(rand_bsr(8192, 8192, 32, 4, np.random.default_rng(0xF32 + f), scale=1 / np.sqrt(128)))
            for f in range(6)]
    X = rng.standard_normal((8192, 512)).astype(np.float32).astype(np.float64)
    F = faust_amd.Faust(facs, dtype=np.float32, act_dtype=np.float32)
    Y = F @ X
    ref = oracle.chain_multiply(facs, X[:, :16], "N")
    rel = np.linalg.norm(Y[:, :16] - ref) / np.linalg.norm(ref)
    assert rel < 1e-5, rel
    Yt = F.T @ X[:, :16]
    reft = oracle.chain_multiply(facs, X[:, :16], "T")
    assert np.linalg.norm(Yt - reft) / np.linalg.norm(reft) < 1e-5
    # the same chain with BF16 activations (config 3a's stated X) is coarser
    Fb = faust_amd.Faust(facs, dtype=np.float32)
    Yb = Fb @ X[:, :16]
    assert np.linalg.norm(Yb - ref) / np.linalg.norm(ref) > 10 * rel


def test_bsr_fp32_handle_values_exact_and_act_dtype_rules(gpu):
    import ctypes as C
    from faust_amd import _lib
    rng = np.random.default_rng(23)
    B = sp.bsr_matrix(rand_bsr(128, 64, 32, 1, rng)).astype(np.float32)
    F = faust_amd.Faust([B], dtype=np.float32)
    L = _lib.lib()
    h = F._facs[0].h
    nb = B.indices.shape[0]
    out = np.empty(nb * 1024, np.float32)
    _lib.check(L.fh_mat_tocpu(h, out.ctypes.data_as(C.c_void_p), None, None))
    want = np.ascontiguousarray(np.transpose(B.data, (0, 2, 1))).ravel()  # column-major blocks
    assert np.array_equal(out, want)
    dt = C.c_int()
    _lib.check(L.fh_faust_get_act_dtype(F.h, C.byref(dt)))
    assert dt.value == _lib.FH_BF16
    Fb = faust_amd.Faust([sp.bsr_matrix(rand_bsr(64, 64, 32, 1, rng))], dtype="bf16")
    assert L.fh_faust_set_act_dtype(Fb.h, _lib.FH_F32) == _lib.FH_ERR_UNSUPPORTED


@pytest.mark.parametrize("dtype", ["bf16", "f32split"])
@pytest.mark.parametrize("op", ["N", "T"])
def test_bsr_autotune_candidates_same_bits(gpu, dtype, op):
    # the BSR kernels the chain autotuner chooses among (5 shared-block, 1 / 2 /
    # 4 ELL tiles per wave, 0 row pointers) give the same bits, and so does the
    # autotuned AUTO product: ragged rows (two slot groups), an empty row, a
    # row wider than the ELL limit (row-pointer fallback), p = 200 / 8 / 512
    from faust_amd import _lib
    rng = np.random.default_rng(31)
    _lib.check(_lib.lib().fh_autotune_clear())
    for wide in (None, 20):
        facs = [_ragged_bsr(rng, r, c, wide_row=wide) for r, c in ((25, 24), (24, 25), (25, 24))]
        if dtype == "bf16":
            facs = [_bsr_rounded(f) for f in facs]
            fdt = "bf16"
        else:
            facs = [sp.bsr_matrix((f.data.astype(np.float32).astype(np.float64), f.indices, f.indptr), shape=f.shape)
                    for f in facs]
            fdt = np.float32
        for p in (200, 8, 512):
            X = bf16_round(rng.standard_normal((32 * (24 if op == "N" else 25), p)))
            outs = {}
            for k in (5, 1, 2, 4, 0, None):
                F = faust_amd.Faust(facs, dtype=fdt)
                if k is not None:
                    F._set_mul_mode(_lib.FH_MUL_BSR_KERNEL(k))
                outs[k] = (F @ X) if op == "N" else (F.T @ X)
            for k, Y in outs.items():
                assert np.array_equal(Y, outs[5]), (k, p, wide)


@pytest.mark.parametrize("op", ["N", "T"])
def test_csr_fp32_autotune_candidates_same_bits(gpu, op):
    # one or two rows per thread, and the autotuned AUTO product: same bits
    from faust_amd import _lib
    rng = np.random.default_rng(32)
    _lib.check(_lib.lib().fh_autotune_clear())
    facs = []
    for s in [(300, 250), (250, 333), (333, 300)]:
        a = rng.standard_normal(s) * (rng.random(s) < rng.uniform(0.0, 0.15, size=(s[0], 1)))
        a[7] = 0.0
        facs.append(sp.csr_matrix(a).astype(np.float32).astype(np.float64))
    for p in (70, 33, 1):
        rows_in = facs[-1].shape[1] if op == "N" else facs[0].shape[0]
        X = rng.standard_normal((rows_in, p)).astype(np.float32).astype(np.float64)
        outs = {}
        for mode in (_lib.FH_MUL_CSR_ROWS1, _lib.FH_MUL_CSR_ROWS2, None):
            F = faust_amd.Faust(facs, dtype=np.float32)
            if mode is not None:
                F._set_mul_mode(mode)
            outs[mode] = (F @ X) if op == "N" else (F.T @ X)
        for mode, Y in outs.items():
            assert np.array_equal(Y, outs[_lib.FH_MUL_CSR_ROWS1]), (mode, p)
