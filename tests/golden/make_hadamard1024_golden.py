"""Golden vectors of the full-size Hadamard hierarchical factorization
(1024 x 1024 Sylvester Hadamard matrix, normalized; 10 factors; splincol
k = 2 factors and k = 512, 256, ..., 2 residuals (ParamsHierarchicalWHT,
factparams.py:1605-1648); right-to-left updates, no packing; 100 PALM4MSA
iterations per local and per global optimisation), written by the oracle
(oracle/, a CPU restatement of faust_hierarchical.hpp and
faust_palm4msa2020.hpp) into tests/golden/hadamard1024_hier.npz:

  traj      every iteration's (call, iteration, relative error, lambda)
  lambda    the final lambda
  sha256    of each final factor, dense FP64 column-major bytes
  row_nnz / col_nnz   nonzeros per row / column of each factor

The run takes about an hour on 8 host cores."""
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hadamard1024_hier.npz")
NLOG2, ITS = 10, 100


def main():
    H = oracle.hadamard(2 ** NLOG2)
    fac, res, kw = oracle.wht_params(NLOG2, ITS)
    t = time.time()
    r = oracle.hierarchical(H, fac, res, record_traj=True, **kw)
    hashes = np.array([hashlib.sha256(np.asfortranarray(f).tobytes(order="F")).hexdigest() for f in r.factors])
    row_nnz = np.array([np.count_nonzero(f, axis=1) for f in r.factors], dtype=np.int32)
    col_nnz = np.array([np.count_nonzero(f, axis=0) for f in r.factors], dtype=np.int32)
    np.savez_compressed(OUT, traj=r.traj, **{"lambda": np.float64(r.lam)}, sha256=hashes, row_nnz=row_nnz,
                        col_nnz=col_nnz, seconds=np.float64(time.time() - t))
    print(OUT, time.time() - t, r.traj[-1])


if __name__ == "__main__":
    main()
