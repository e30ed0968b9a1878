"""faust_amd -- MI355X (gfx950) backend of the FAuST hot paths.

pyfaust-shaped Python layer over libfaust_hip.so:
  Faust(factors) @ X, F.T @ X, F.toarray(), F.norm(2)       (Faust x dense)
  fact.palm4msa / fact.hierarchical (backend 2020, on GPU)   (PALM4MSA loop)
  proj.sp / spcol / splin / splincol / supp                  (projectors)
  rand(...), wht(n)                                           (constructors)
"""
import numpy as np
import scipy.sparse as sp

from . import fact, factparams, proj
from ._lib import FaustHipError
from .butterfly import bitrev_perm
from .faust import Factor, Faust

__all__ = ["Faust", "Factor", "rand", "rand_bsr", "rand_butterfly", "bitrev_perm", "wht", "fact", "factparams",
           "proj", "FaustHipError"]


def rand_csr(nrows, ncols, nnz_per_row, rng, values="normal", dtype=np.float64):
    """CSR with exactly nnz_per_row distinct uniformly drawn columns per row
    (MatSparse::randMat per_row structure, faust_MatSparse.hpp:1280-1345)."""
    k = int(nnz_per_row)
    cols = np.argsort(rng.random((nrows, ncols)), axis=1)[:, :k] if ncols <= 4096 else \
        np.stack([rng.choice(ncols, size=k, replace=False) for _ in range(nrows)])
    cols = np.sort(cols, axis=1)
    vals = rng.standard_normal((nrows, k)) if values == "normal" else rng.random((nrows, k))
    indptr = np.arange(0, nrows * k + 1, k, dtype=np.int32)
    return sp.csr_matrix((vals.ravel().astype(dtype), cols.ravel().astype(np.int32), indptr),
                         shape=(nrows, ncols))


def rand(num_rows, num_cols, num_factors=None, dim_sizes=None, density=None, fac_type="sparse", per_row=True,
         seed=0, dtype=np.float64, values="uniform"):
    """pyfaust.rand (pyfaust.py:4597-4808) for square/rectangular chains of
    'sparse' or 'dense' factors.  density*ncols nonzeros per row; values U(0,1)
    like MatSparse::randMat (values='normal' draws N(0,1) instead)."""
    rng = np.random.default_rng(seed)
    if num_factors is None:
        num_factors = 5
    if isinstance(num_factors, (list, tuple)):
        num_factors = int(rng.integers(num_factors[0], num_factors[1] + 1))
    if density is None:
        density = 5.0 / num_cols
    if dim_sizes is None:
        dims = [num_rows] + [num_cols] * num_factors
    else:
        lo, hi = (dim_sizes, dim_sizes) if np.isscalar(dim_sizes) else dim_sizes
        dims = [num_rows] + [int(rng.integers(lo, hi + 1)) for _ in range(num_factors - 1)] + [num_cols]
    facs = []
    for i in range(num_factors):
        r, c = dims[i], dims[i + 1]
        if fac_type == "dense":
            a = rng.standard_normal((r, c)) if values == "normal" else rng.random((r, c))
            facs.append(np.asfortranarray(a.astype(dtype)))
        else:
            k = max(1, int(density * c)) if per_row else max(1, int(density * r))
            if per_row:
                facs.append(rand_csr(r, c, k, rng, values, dtype))
            else:
                facs.append(sp.csr_matrix(rand_csr(c, r, k, rng, values, dtype).T))
    return Faust(facs, dtype=dtype)


def rand_bsr(num_rows, num_cols, bnrows, bncols, num_factors=None, density=.1, dev="gpu", dtype="float64", seed=0):
    """pyfaust.rand_bsr (pyfaust.py:4484-4560; TransformHelper::randBSRFaust,
    faust_TransformHelper.hpp:1430-1449): a square Faust of random BSR
    factors, each with round(density * num_rows*num_cols/(bnrows*bncols))
    nonzero blocks of U(0, 1) values.  num_factors: an int, a (min, max)
    pair (inclusive, drawn uniformly) or None for 5.  dtype 'float32' or
    'bf16' keeps the factors BSR on the GPU (32x32 blocks go to MFMA, other
    sizes run as CSR); 'float64' runs them as FP64 CSR factors of the same
    values (the GPU stores BSR blocks in BF16 or FP32 only)."""
    from .bsr import rand_bsr_factor
    if num_factors is None:
        lo = hi = 5
    elif isinstance(num_factors, (int, np.integer)):
        lo = hi = int(num_factors)
    elif isinstance(num_factors, tuple) and len(num_factors) == 2:
        lo, hi = int(num_factors[0]), int(num_factors[1])
    else:
        raise ValueError("num_factors must be None, a int or a tuple of int")
    if num_rows != num_cols:
        raise ValueError("currently only random square BSR Fausts can be generated.")
    if num_rows % bnrows != 0 or num_cols % bncols != 0:
        raise ValueError("the size of blocks must evenly divide the size of Faust matrices")
    if not dev.startswith("gpu"):
        raise ValueError("this build runs Fausts on the GPU only (dev='gpu')")
    bnnz = int(round(num_rows * num_cols / bnrows / bncols * density))
    if bnnz == 0:
        raise ValueError("the nonzero blocks are too large for this Faust/matrix size.")
    rng = np.random.default_rng(seed)
    J = int(rng.integers(lo, hi + 1))
    facs = [rand_bsr_factor(num_rows, num_cols, bnrows, bncols, bnnz, rng) for _ in range(J)]
    if dtype in ("float64", np.float64):
        return Faust([sp.csr_matrix(f) for f in facs], dtype=np.float64)
    if dtype in ("float32", np.float32):
        return Faust(facs, dtype=np.float32)
    if dtype == "bf16":
        return Faust(facs, dtype="bf16")
    raise ValueError("dtype must be 'float64', 'float32' or 'bf16'")


def wht_factors(n):
    """The log2(n) butterfly factors of the normalised Walsh-Hadamard transform
    (faust_WHT.hpp:40-104 structure: each factor B*P with entries +-1/sqrt(2))."""
    L = int(np.log2(n))
    if 2 ** L != n:
        raise ValueError("n must be a power of two")
    h = n // 2
    rows = np.repeat(np.arange(n), 2)
    cols = np.empty(2 * n, dtype=np.int64)
    vals = np.empty(2 * n)
    for i in range(h):
        cols[4 * i:4 * i + 2] = [2 * i, 2 * i + 1]
        vals[4 * i:4 * i + 2] = [1, 1]
        cols[4 * i + 2:4 * i + 4] = [2 * i, 2 * i + 1]
        vals[4 * i + 2:4 * i + 4] = [1, -1]
    # rows i and i+h: apply the perfect shuffle so that row r of the factor combines
    # entries 2*(r mod h) and 2*(r mod h)+1 with the sign of (r >= h)
    data, indices, indptr = [], [], [0]
    for r in range(n):
        i = r % h
        s = -1.0 if r >= h else 1.0
        indices += [2 * i, 2 * i + 1]
        data += [1.0 / np.sqrt(2), s / np.sqrt(2)]
        indptr.append(len(indices))
    B = sp.csr_matrix((np.array(data), np.array(indices), np.array(indptr)), shape=(n, n))
    return [B.copy() for _ in range(L)]


def wht(n, normed=True):
    """Faust of the Walsh-Hadamard transform (pyfaust.wht, pyfaust.py:3717)."""
    facs = wht_factors(n)
    if not normed:
        facs[0] = facs[0] * np.sqrt(n)
    return Faust(facs)


def rand_butterfly(n, dtype="float64", dev="gpu", diag_opt=False):
    """pyfaust.rand_butterfly (pyfaust.py:4810-4848): the product of log2(n)
    factors on the butterfly supports with standard normal values drawn with
    np.random.randn.  Real dtypes only; diag_opt is not provided."""
    from .butterfly import rand_butterfly_factors
    if diag_opt:
        raise NotImplementedError("diag_opt (opt_butterfly_faust) is not provided")
    if dtype not in ("float64", "float32", np.float64, np.float32):
        raise ValueError("dtype must be 'float64' or 'float32'")
    dt = np.dtype(dtype).type
    return Faust(rand_butterfly_factors(n, dt), dtype=dt)
