"""Batched PALM4MSA engine vs the oracle's palm4msa2 (same schedule:
left-to-right, packed products, sp projector).

In FP64 with cold power-iteration starts the engine differs from the oracle
only in reduction orders (power iteration, lambda), so supports must be
identical and factors/lambda agree to 1e-9 relative after 15 iterations on
tie-free random data.  The FP32 + FP64-refinement path is checked through
its FP64-verified relative error."""
import numpy as np
import pytest

import faust_amd
import oracle
from faust_amd.batched import palm4msa_batched

pytestmark = pytest.mark.gpu


def _targets(seed, sizes, J=3, nnz=4, noise=1e-6):
    rng = np.random.default_rng(seed)
    import faust_amd
    out = []
    for n in sizes:
        P = np.eye(n)
        for _ in range(J):
            P = P @ faust_amd.rand_csr(n, n, nnz, rng).toarray()
        out.append(np.asfortranarray(P + noise * rng.standard_normal((n, n))))
    return out


def _oracle_run(A, n, its, J=3):
    cons = [("sp", 4 * n, n, n)] * J
    return oracle.palm4msa(A, cons, num_its=its)


def test_batched_fp64_matches_oracle(gpu):
    sizes = [32, 64, 48, 64]
    As = _targets(1, sizes)
    its = 15
    res = palm4msa_batched(As, 3, k=[4 * n for n in sizes], dtype="f64", num_its=its, refine_its=0,
                           warm_power=False, autotune=False)
    for b, (A, n) in enumerate(zip(As, sizes)):
        ref = _oracle_run(A, n, its)
        assert abs(res.lambdas[b] - ref.lam) <= 1e-9 * abs(ref.lam)
        for f in range(3):
            got, exp = res.factors[b][f], ref.factors[f]
            assert np.array_equal(got != 0, exp != 0), f"support differs (target {b}, factor {f})"
            assert np.linalg.norm(got - exp) <= 1e-9 * np.linalg.norm(exp)
        P = np.eye(n)
        for f in ref.factors:
            P = P @ f
        rel = np.linalg.norm(ref.lam * P - A) / np.linalg.norm(A)
        assert abs(res.relerr[b] - rel) <= 1e-9


def test_batched_graph_equals_eager(gpu):
    sizes = [64, 128, 64]
    As = _targets(2, sizes)
    kw = dict(k=[4 * n for n in sizes], dtype="f32", num_its=10, refine_its=3, autotune=False)
    a = palm4msa_batched(As, 3, use_graph=True, **kw)
    b = palm4msa_batched(As, 3, use_graph=False, **kw)
    assert np.array_equal(a.relerr, b.relerr)
    assert np.array_equal(a.lambdas, b.lambdas)


def test_batched_mixed_precision_reaches_fp64_error(gpu):
    sizes = [64, 128, 256, 64, 128, 512]
    As = _targets(3, sizes)
    its = 60
    mixed = palm4msa_batched(As, 3, k=[4 * n for n in sizes], dtype="f32", num_its=its, refine_its=10)
    plain = palm4msa_batched(As, 3, k=[4 * n for n in sizes], dtype="f64", num_its=its + 10, refine_its=0,
                             warm_power=False)
    assert np.all(np.isfinite(mixed.relerr))
    # the frozen-support FP64 sweeps bring the FP32 run to the FP64 run's accuracy
    assert np.median(mixed.relerr / plain.relerr) < 1.5
    assert np.all(mixed.relerr < 1.0)


def test_batched_supp_masks(gpu):
    rng = np.random.default_rng(5)
    n = 64
    masks = [(rng.random((n, n)) < 0.1).astype(float) + np.eye(n) for _ in range(2)]
    masks = [np.minimum(m, 1.0) for m in masks]
    A = np.asfortranarray((masks[0] * rng.standard_normal((n, n))) @ (masks[1] * rng.standard_normal((n, n))))
    res = palm4msa_batched([A], 2, supports=[masks], dtype="f64", num_its=30, refine_its=0, warm_power=False,
                           autotune=False)
    for f in range(2):
        assert np.all(res.factors[0][f][masks[f] == 0] == 0)
    ref = oracle.palm4msa(A, [("supp", 0, n, n, True, False, masks[0]), ("supp", 0, n, n, True, False, masks[1])],
                          num_its=30)
    assert abs(res.lambdas[0] - ref.lam) <= 1e-8 * abs(ref.lam)


@pytest.mark.parametrize("tile", ["32", "64", "65"])
def test_batched_gemm_tiles_agree(gpu, monkeypatch, tile):
    # the 32/64 VALU tiles and the 64x64 FP32 MFMA tile run the same FP32
    # iterations (sizes that are not tile multiples included); compared with
    # the FP64 batched run of the same problem
    monkeypatch.setenv("FH_BGEMM_TILE", tile)
    rng = np.random.default_rng(9)
    As, ks = [], []
    for n in (64, 100, 160):
        P = np.eye(n)
        for _ in range(3):
            P = P @ (faust_amd.rand_csr(n, n, 4, rng).toarray())
        As.append(np.asfortranarray(P))
        ks.append(4 * n)
    r32 = palm4msa_batched(As, 3, k=ks, dtype="f32", num_its=30, refine_its=0, autotune=True)
    r64 = palm4msa_batched(As, 3, k=ks, dtype="f64", num_its=30, refine_its=0, autotune=False)
    assert np.all(np.isfinite(r32.relerr))
    assert np.all(np.abs(r32.relerr - r64.relerr) <= 0.02 * r64.relerr + 1e-5)


def test_batched_fp32_sparse_step_sizes_track_dense(gpu, monkeypatch):
    # the FP32 engine's step-size power iterations on the compressed sparse
    # factors (default for sp projections keeping at most 1/8 of a factor)
    # against the same iterations on the dense partial products: the two
    # differ only by the FP32 rounding of the dense products, so after 60
    # iterations the FP64-verified errors agree closely and every target
    # converges the same way; sizes with rows of up to 2048 / 256 threads,
    # and 300 (sparse x dense blocks of 6 columns over two rows per thread)
    sizes = [64, 128, 96, 256, 64, 512, 300]
    As = _targets(5, sizes)
    ks = [4 * n for n in sizes]
    rel = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("FH_BATCH_SPARSE_POWER", mode)
        res = palm4msa_batched(As, 3, k=ks, dtype="f32", num_its=60, refine_its=0, return_factors=False)
        rel[mode] = np.asarray(res.relerr)
    assert np.all(np.isfinite(rel["1"])) and np.all(rel["1"] < 1.0)
    assert np.all(np.abs(rel["1"] - rel["0"]) <= 1e-3 * rel["0"] + 1e-6), (rel["1"], rel["0"])


@pytest.mark.parametrize("sizes", [[64, 96, 256, 512, 200], [512, 1024, 64]])
def test_batched_sparse_power_register_rows_same_bits(gpu, monkeypatch, sizes):
    # edges <= 512: the sparse power iterations keep each thread's rows in
    # registers (short rows padded with -0 x +0 terms); FH_BATCH_POWER_REG=0
    # reads them from memory every round.  The two run the same fma chains,
    # so factors, lambdas and errors must agree bit for bit.  With an edge
    # above 512 both settings take the memory path.
    As = _targets(11, sizes)
    ks = [4 * n for n in sizes]
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("FH_BATCH_POWER_REG", mode)
        out[mode] = palm4msa_batched(As, 3, k=ks, dtype="f32", num_its=30, refine_its=0, autotune=False)
    a, b = out["1"], out["0"]
    assert np.array_equal(np.asarray(a.relerr), np.asarray(b.relerr))
    assert np.array_equal(np.asarray(a.lambdas), np.asarray(b.lambdas))
    assert np.all(np.asarray(a.relerr) < 1.0)
    for t in range(len(sizes)):
        for f in range(3):
            assert np.array_equal(a.factors[t][f], b.factors[t][f]), (t, f)
