"""Writes the golden vectors of tests/golden/ from the CPU oracle.

The reference itself cannot run here (Eigen 3.4 is absent, pyfaust is not
installed), so these fixtures pin the oracle's own outputs: later changes to
the oracle must reproduce them bit for bit (tests/test_golden.py), and the
GPU tests compare the HIP path to the same values.

python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def chain_case():
    rng = np.random.default_rng(20240601)
    shapes = [(24, 16), (16, 32), (32, 32), (32, 20)]
    facs = []
    for i, (r, c) in enumerate(shapes):
        a = rng.standard_normal((r, c)) * (rng.random((r, c)) < 0.3)
        facs.append(sp.csr_matrix(a) if i != 2 else a)
    X = rng.standard_normal((20, 6))
    X2 = rng.standard_normal((24, 6))
    d = {"X": X, "X2": X2, "YN": oracle.chain_multiply(facs, X, "N"), "YT": oracle.chain_multiply(facs, X2, "T"),
         "P": oracle.get_product(facs), "norm2": np.array(oracle.spectral_norm(facs)[0])}
    for i, f in enumerate(facs):
        d[f"F{i}"] = f.toarray() if sp.issparse(f) else f
        d[f"F{i}_sparse"] = np.array(sp.issparse(f))
    np.savez(os.path.join(HERE, "chain_small.npz"), **d)


def prox_case():
    rng = np.random.default_rng(7)
    M = rng.integers(-2, 3, (16, 12)).astype(float)  # heavy ties
    d = {"M": M}
    for name, k in (("sp", 30), ("spcol", 3), ("splin", 3), ("splincol", 2)):
        d[name] = oracle.prox(name, M, k)
    np.savez(os.path.join(HERE, "prox_ties.npz"), **d)


def hadamard_case():
    H = oracle.hadamard(32)
    fac, res, kw = oracle.wht_params(5, 30)
    r = oracle.hierarchical(H, fac, res, record_traj=True, **kw)
    d = {"lambda": np.array(r.lam), "traj": r.traj}
    for i, f in enumerate(r.factors):
        d[f"S{i}"] = f
    np.savez(os.path.join(HERE, "hadamard32_hier.npz"), **d)


def x0_case():
    np.savez(os.path.join(HERE, "power_x0.npz"), x0_1024=oracle.power_x0(1024))


if __name__ == "__main__":
    chain_case()
    prox_case()
    hadamard_case()
    x0_case()
    print("golden vectors written to", HERE)
