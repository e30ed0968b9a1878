"""The batched PALM4MSA workload at full size: 512 square targets, n drawn
from {64, 128, 256, 512}, each a product of 3 sparse factors (4 entries per
row) plus N(0, 1e-6) noise, factorized into 3 factors with sp k = 4n.

* the batched FP64 form (220 iterations, the oracle's cold power-iteration
  starts) gives a finite error on every target, and on sampled targets the
  oracle's PALM4MSA error (faust_palm4msa2020.hpp) to 1e-9;
* the tuned form (FP32 iterations + FP64 frozen-support sweeps, checked
  every 10 iterations) reaches, FP64-verified, that error on every target.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_batched_512_targets_time_to_fp64_error(gpu):
    import bench
    from faust_amd.batched import palm4msa_batched
    print("batched PALM4MSA, 512 targets, FP64 220 its then FP32 <= 400 + FP64 20 checked every 10", flush=True)
    sizes, As = bench.palm_targets()
    ks = [4 * int(n) for n in sizes]
    # the FP64 form with the oracle's cold power-iteration starts
    b64 = palm4msa_batched(As, 3, k=ks, dtype="f64", num_its=220, refine_its=0, warm_power=False,
                           return_factors=False)
    assert np.all(np.isfinite(b64.relerr)) and np.all(b64.relerr < 1.0)
    # sampled targets against the oracle's PALM4MSA (220 iterations): only
    # reduction orders differ
    for n in (64, 128):
        for i in np.flatnonzero(sizes == n)[:2]:
            ref = oracle.palm4msa(As[i], [("sp", 4 * n, n, n)] * 3, num_its=220)
            P = ref.lam * ref.factors[0] @ ref.factors[1] @ ref.factors[2]
            e = np.linalg.norm(P - As[i]) / np.linalg.norm(As[i])
            assert abs(b64.relerr[i] - e) <= 1e-9 * e, (n, i)
    goal = b64.relerr * (1 + 1e-9)
    tuned = palm4msa_batched(As, 3, k=ks, dtype="f32", num_its=400, refine_its=20, return_factors=False,
                             check_every=10, target_err=goal)
    assert np.all(np.isfinite(tuned.relerr))
    assert np.all(tuned.iters_to_target >= 0), "a target never reached the FP64 form's error"
    assert tuned.times_ms[4] >= 0
    assert np.all(tuned.relerr <= goal)
