"""The oracle reproduces the committed golden vectors bit for bit
(tests/golden/make_golden.py wrote them)."""
import os

import numpy as np
import scipy.sparse as sp

import oracle

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _facs(d):
    out = []
    i = 0
    while f"F{i}" in d:
        f = d[f"F{i}"]
        out.append(sp.csr_matrix(f) if bool(d[f"F{i}_sparse"]) else f)
        i += 1
    return out


def test_chain_golden():
    d = np.load(os.path.join(G, "chain_small.npz"))
    facs = _facs(d)
    assert np.array_equal(oracle.chain_multiply(facs, d["X"], "N"), d["YN"])
    assert np.array_equal(oracle.chain_multiply(facs, d["X2"], "T"), d["YT"])
    assert np.array_equal(oracle.get_product(facs), d["P"])
    assert oracle.spectral_norm(facs)[0] == float(d["norm2"])


def test_prox_golden():
    d = np.load(os.path.join(G, "prox_ties.npz"))
    for name, k in (("sp", 30), ("spcol", 3), ("splin", 3), ("splincol", 2)):
        assert np.array_equal(oracle.prox(name, d["M"], k), d[name])


def test_hadamard_golden():
    d = np.load(os.path.join(G, "hadamard32_hier.npz"))
    fac, res, kw = oracle.wht_params(5, 30)
    r = oracle.hierarchical(oracle.hadamard(32), fac, res, record_traj=True, **kw)
    assert r.lam == float(d["lambda"])
    assert np.array_equal(r.traj, d["traj"])
    for i, f in enumerate(r.factors):
        assert np.array_equal(f, d[f"S{i}"])


def test_power_x0_golden():
    d = np.load(os.path.join(G, "power_x0.npz"))
    assert np.array_equal(oracle.power_x0(1024), d["x0_1024"])
