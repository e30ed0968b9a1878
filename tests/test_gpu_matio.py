"""Faust.save / Faust.load / Faust(filepath) through the GPU handles, and
fh_mat_tocpu for CSR and BSR handles."""
import numpy as np
import pytest
import scipy.sparse as sp

import faust_amd
from faust_amd.bsr import bf16_bits_to_f32, f32_to_bf16_bits, rand_bsr

pytestmark = pytest.mark.gpu


def _dense(m):
    return m.toarray() if sp.issparse(m) else np.asarray(m)


def test_fp64_faust_round_trip(gpu, tmp_path):
    rng = np.random.default_rng(2)
    facs = [faust_amd.rand_csr(50, 40, 5, rng), rng.standard_normal((40, 30)), faust_amd.rand_csr(30, 60, 4, rng)]
    F = faust_amd.Faust(facs)
    path = str(tmp_path / "F.mat")
    F.save(path)
    for G in (faust_amd.Faust(path), faust_amd.Faust(filepath=path), faust_amd.Faust.load(path)):
        assert G.shape == F.shape and G.numfactors() == 3
        for i in range(3):
            assert np.array_equal(G.factors(i), F.factors(i))
        X = rng.standard_normal((60, 7))
        assert np.array_equal(G @ X, F @ X)
    # the transposed Faust is written as transposed factors in reverse order
    F.T.save(path)
    Gt = faust_amd.Faust(path)
    assert Gt.shape == F.T.shape
    for i in range(3):
        assert np.array_equal(Gt.factors(i), F.factors(2 - i).T)
    # the dense products associate the factors in different orders
    A, B = Gt.toarray(), F.T.toarray()
    assert np.linalg.norm(A - B) <= 1e-13 * np.linalg.norm(B)


def test_bsr_fausts_round_trip(gpu, tmp_path):
    rng = np.random.default_rng(3)
    facs = [rand_bsr(256, 256, 32, 2, np.random.default_rng(30 + f), scale=1 / np.sqrt(64)) for f in range(3)]
    X = rng.standard_normal((256, 40))
    # FP32 blocks: the file holds the exact FP32 values
    F32 = faust_amd.Faust(facs, dtype=np.float32)
    p32 = str(tmp_path / "b32.mat")
    F32.save(p32)
    G32 = faust_amd.Faust(p32)
    for i in range(3):
        assert np.array_equal(G32._facs[i].to_host().toarray(), facs[i].toarray().astype(np.float32))
    assert np.array_equal(G32 @ X, F32 @ X)
    # BF16 blocks: the file holds the stored BF16 values (as single)
    Fb = faust_amd.Faust(facs, dtype="bf16")
    pb = str(tmp_path / "bb.mat")
    Fb.save(pb)
    Gb = faust_amd.Faust(pb)
    for i in range(3):
        want = bf16_bits_to_f32(f32_to_bf16_bits(facs[i].toarray()))
        assert np.array_equal(Gb._facs[i].to_host().toarray(), want)
    # loaded as FP32 hi/lo blocks whose lo halves are zero: the same product
    assert np.array_equal(Gb @ X, Fb @ X)


def test_tocpu_handles(gpu):
    rng = np.random.default_rng(4)
    # FP32 CSR
    A = sp.csr_matrix(rng.standard_normal((33, 17)) * (rng.random((33, 17)) < 0.3)).astype(np.float32)
    f = faust_amd.Factor(A, np.float32)
    H = f.to_host()
    assert H.dtype == np.float32 and np.array_equal(H.toarray(), A.toarray())
    # FP32 BSR without the host copy: hi + lo of the stored pair
    B = rand_bsr(128, 96, 32, 2, rng)
    fb = faust_amd.Factor(B, np.float32)
    fb._host = None
    Hb = fb.to_host()
    assert Hb.blocksize == (32, 32)
    ref = B.toarray().astype(np.float32)
    assert np.all(np.abs(Hb.toarray() - ref) <= 2.0 ** -16 * np.abs(ref))
    assert np.array_equal(Hb.toarray() != 0, ref != 0)


def test_rand_bsr(gpu, tmp_path):
    # pyfaust.rand_bsr(100, 100, 20, 10, num_factors=6): 6 factors of 5 blocks
    F = faust_amd.rand_bsr(100, 100, 20, 10, num_factors=6, seed=1)
    assert F.numfactors() == 6 and F.shape == (100, 100)
    assert F.nnz_sum() == 6 * 5 * 200
    G = faust_amd.rand_bsr(256, 256, 32, 32, num_factors=(2, 4), density=0.25, dtype="float32", seed=2)
    assert 2 <= G.numfactors() <= 4
    P = np.eye(256)
    for i in range(G.numfactors()):
        P = P @ G._facs[i].to_host().toarray().astype(np.float64)
    X = np.random.default_rng(3).standard_normal((256, 8))
    Y = G @ X
    assert np.linalg.norm(Y - P @ X) / np.linalg.norm(P @ X) < 1e-2  # BF16 activations
    path = str(tmp_path / "r.mat")
    G.save(path)
    assert np.array_equal(faust_amd.Faust(path) @ X, Y)
    with pytest.raises(ValueError):
        faust_amd.rand_bsr(100, 80, 20, 10)
    with pytest.raises(ValueError):
        faust_amd.rand_bsr(100, 100, 30, 10)
