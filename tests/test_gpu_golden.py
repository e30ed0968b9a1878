"""The HIP path reproduces the committed golden vectors bit for bit."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import faust_amd
from faust_amd import fact, factparams, proj

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_chain_golden_gpu(gpu):
    d = np.load(os.path.join(G, "chain_small.npz"))
    facs = []
    i = 0
    while f"F{i}" in d:
        f = d[f"F{i}"]
        facs.append(sp.csr_matrix(f) if bool(d[f"F{i}_sparse"]) else f)
        i += 1
    F = faust_amd.Faust(facs)
    assert np.array_equal(F @ d["X"], d["YN"])
    assert np.array_equal(F.T @ d["X2"], d["YT"])
    assert np.array_equal(F.toarray(), d["P"])
    assert F.norm(2) == float(d["norm2"])


def test_prox_golden_gpu(gpu):
    d = np.load(os.path.join(G, "prox_ties.npz"))
    M = d["M"]
    for name, k in (("sp", 30), ("spcol", 3), ("splin", 3), ("splincol", 2)):
        assert np.array_equal(getattr(proj, name)(M.shape, k)(M), d[name])


def test_hadamard_golden_gpu(gpu):
    d = np.load(os.path.join(G, "hadamard32_hier.npz"))
    import oracle
    res = fact.hierarchical(oracle.hadamard(32), factparams.ParamsHierarchicalWHT(5, num_its=30), ret_raw=True)
    assert res.lam == float(d["lambda"])
    assert np.array_equal(res.traj, d["traj"])
    for i, f in enumerate(res.dense_factors()):
        assert np.array_equal(f, d[f"S{i}"])
