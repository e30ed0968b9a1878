"""Config 2 at full size: hierarchical PALM4MSA of the 1024 x 1024 Sylvester
Hadamard matrix into 10 factors (ParamsHierarchicalWHT, factparams.py:
1605-1648: splincol k = 2 factors, k = 512 ... 2 residuals, right-to-left
updates, no packing; 100 iterations per local and per global PALM4MSA,
faust_hierarchical.hpp:1-150) on the GPU, compared bit for bit with the
oracle's run committed in tests/golden/hadamard1024_hier.npz (written by
tests/golden/make_hadamard1024_golden.py, 44 minutes on 8 host cores): every
iteration's relative error and lambda, the final lambda and every factor's
bytes (sha256).  Also the properties that need no oracle: exactly 2 nonzeros
per row and per column in every factor, relative error below 1e-12, and
lambda * S_1 ... S_10 = H."""
import hashlib
import os
import time

import numpy as np
import pytest

import oracle
from faust_amd import fact, factparams

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_hierarchical_hadamard_1024_full_size(gpu):
    d = np.load(os.path.join(G, "hadamard1024_hier.npz"))
    H = oracle.hadamard(1024)
    print("hierarchical Hadamard 1024, 10 factors, 100 its per PALM4MSA", flush=True)
    t0 = time.perf_counter()
    res = fact.hierarchical(H, factparams.ParamsHierarchicalWHT(10, num_its=100), ret_raw=True)
    print(f"GPU run {time.perf_counter() - t0:.1f} s", flush=True)
    facs = res.dense_factors()
    assert len(facs) == 10
    # bit for bit against the oracle's run
    assert res.traj.shape == d["traj"].shape
    assert np.array_equal(res.traj, d["traj"])
    assert res.lam == float(d["lambda"])
    for i, f in enumerate(facs):
        assert hashlib.sha256(np.asfortranarray(f).tobytes(order="F")).hexdigest() == str(d["sha256"][i]), i
    # supports: two nonzeros per row and per column in every factor
    for i, f in enumerate(facs):
        assert np.array_equal(np.count_nonzero(f, axis=1), d["row_nnz"][i])
        assert np.array_equal(np.count_nonzero(f, axis=0), d["col_nnz"][i])
        assert (np.count_nonzero(f, axis=1) == 2).all() and (np.count_nonzero(f, axis=0) == 2).all()
    # the factorization is exact
    P = facs[0]
    for f in facs[1:]:
        P = P @ f
    err = np.linalg.norm(res.lam * P - H) / np.linalg.norm(H)
    assert err < 1e-12
    assert res.traj[-1, 2] < 1e-12
