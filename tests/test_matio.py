"""Faust .mat files on the host: the reference layout (faust_factors cell;
BSR as the 5-cell of faust_MatBSR.hpp:387-480) read back by the reference's
own reader, restated here from pyfaust.py:2620-2635."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.io import loadmat

from faust_amd import matio
from faust_amd.bsr import rand_bsr


def _dense(m):
    return m.toarray() if sp.issparse(m) else np.asarray(m)


def _reference_reader(path):
    # pyfaust Faust.load: cells as returned by loadmat, BSR cells rebuilt
    # from their column-major blocks.  The isinstance guard keeps the check
    # to cell entries: scipy >= 1.15 returns sparse cells as coo_array, which
    # the reference's unguarded factors[i][0][0] cannot index.
    factors = loadmat(path)["faust_factors"][0].tolist()
    for i in range(len(factors)):
        is_cell = isinstance(factors[i], np.ndarray) and factors[i].dtype == object
        if is_cell and factors[i][0][0].dtype == "<U3" and factors[i][0][0][0] == "bsr":
            nrows, ncols, bnnz = factors[i][0][1][0]
            bcolinds = factors[i][0][2][0]
            browptr = factors[i][0][3][0]
            bdata = factors[i][0][4][0]
            bnrows = int(nrows / (browptr.shape[0] - 1))
            bncols = int(bdata.shape[0] / bnrows / bnnz)
            bdata_ = np.empty((bnnz, bnrows, bncols))
            for bi in range(bnnz):
                bdata_[bi] = bdata.reshape(bnnz, bncols, bnrows)[bi].T
            factors[i] = sp.bsr_matrix((bdata_, bcolinds, browptr), shape=(nrows, ncols))
    return factors


def test_round_trip_and_reference_reader(tmp_path):
    rng = np.random.default_rng(0)
    facs = [sp.random(20, 30, density=0.2, random_state=1, format="csr"),
            rng.standard_normal((30, 64)),
            rand_bsr(64, 96, 32, 2, rng),
            sp.bsr_matrix(rng.standard_normal((6, 8)) * (rng.random((6, 8)) < 0.5), blocksize=(2, 4))]
    path = str(tmp_path / "F.mat")
    matio.write_factors(path, facs)
    back = matio.read_factors(path)
    assert [type(b).__name__ for b in back] == ["csr_matrix", "ndarray", "bsr_matrix", "bsr_matrix"]
    for a, b in zip(facs, back):
        assert np.array_equal(_dense(a), _dense(b))
    assert back[2].blocksize == (32, 32) and back[3].blocksize == (2, 4)
    ref = _reference_reader(path)
    for a, b in zip(facs, ref):
        assert np.array_equal(_dense(a), _dense(b.toarray() if sp.issparse(b) else b))
    raw = loadmat(path)["faust_factors"]
    assert raw.shape == (1, 4)
    assert sp.issparse(raw[0, 0]) and raw[0, 0].dtype == np.float64
    params = raw[0, 2][0, 1]
    assert params.dtype == np.int32 and params.tolist() == [[64, 96, 4]]


def test_single_precision_and_errors(tmp_path):
    rng = np.random.default_rng(1)
    B = rand_bsr(64, 64, 32, 1, rng)
    B32 = sp.bsr_matrix((B.data.astype(np.float32), B.indices, B.indptr), shape=B.shape)
    path = str(tmp_path / "s.mat")
    matio.write_factors(path, [B32, rng.standard_normal((64, 3)).astype(np.float32)])
    back = matio.read_factors(path)
    assert back[0].data.dtype == np.float32 and back[1].dtype == np.float32
    assert np.array_equal(back[0].toarray(), B32.toarray())
    with pytest.raises(ValueError):
        matio.write_factors(path, [])
    from scipy.io import savemat
    savemat(str(tmp_path / "x.mat"), {"other": np.eye(2)})
    with pytest.raises(ValueError):
        matio.read_factors(str(tmp_path / "x.mat"))


def test_rand_bsr_factor_structure():
    # BSRMat::rand: bnnz distinct block positions, U(0,1) values, capped at
    # the number of blocks
    from faust_amd.bsr import rand_bsr_factor
    rng = np.random.default_rng(5)
    B = rand_bsr_factor(100, 100, 20, 10, 13, rng)
    assert B.blocksize == (20, 10) and B.indices.shape[0] == 13
    pos = set()
    for br in range(5):
        for b in range(B.indptr[br], B.indptr[br + 1]):
            pos.add((br, int(B.indices[b])))
    assert len(pos) == 13
    assert np.all(B.data >= 0) and np.all(B.data < 1)
    assert rand_bsr_factor(64, 64, 32, 32, 99, rng).indices.shape[0] == 4
