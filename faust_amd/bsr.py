"""BSR factor upload (MatBSR: faust_MatBSR.h:120-150, blocks column-major)."""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import check, lib


def f32_to_bf16_bits(a):
    """Round-to-nearest-even float32 -> bfloat16 bit patterns (uint16)."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r


def bf16_bits_to_f32(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def bsr_handle(M, dt):
    """Upload a scipy BSR matrix as a device BSR handle (FH_BF16 or FH_F32 values)."""
    B = sp.bsr_matrix(M)
    B.sort_indices()
    R, Cb = B.blocksize
    data = np.ascontiguousarray(np.transpose(B.data, (0, 2, 1)))  # column-major blocks
    if dt == _lib.FH_BF16:
        vals = f32_to_bf16_bits(data)
    elif dt == _lib.FH_F32:
        vals = np.ascontiguousarray(data, dtype=np.float32)
    else:
        raise ValueError("BSR factors are stored BF16 or FP32 on the GPU")
    bp = np.ascontiguousarray(B.indptr, dtype=np.int32)
    bc = np.ascontiguousarray(B.indices, dtype=np.int32)
    h = C.c_void_p()
    check(lib().fh_mat_bsr_create(dt, B.shape[0], B.shape[1], R, Cb, len(bc), bp.ctypes.data_as(C.c_void_p),
                                  bc.ctypes.data_as(C.c_void_p), vals.ctypes.data_as(C.c_void_p), 0, C.byref(h)))
    return h


def rand_bsr(nrows, ncols, bs, blocks_per_row, rng, scale=1.0):
    """BSR with exactly blocks_per_row distinct uniformly drawn block columns per
    block-row and N(0, scale^2) values."""
    nbr, nbc = nrows // bs, ncols // bs
    cols = np.sort(np.argsort(rng.random((nbr, nbc)), axis=1)[:, :blocks_per_row], axis=1)
    data = rng.standard_normal((nbr * blocks_per_row, bs, bs)) * scale
    indptr = np.arange(0, nbr * blocks_per_row + 1, blocks_per_row)
    return sp.bsr_matrix((data, cols.ravel(), indptr), shape=(nrows, ncols))


def rand_bsr_factor(nrows, ncols, bm, bn, bnnz, rng):
    """BSRMat::rand (faust_MatBSR.hpp:693-760): bnnz blocks (at most all of
    them) at distinct block positions drawn uniformly without replacement,
    values U(0, 1); blocks sorted by (block row, block column)."""
    if nrows % bm or ncols % bn:
        raise ValueError("the size of blocks must evenly divide the size of the matrix")
    nbr, nbc = nrows // bm, ncols // bn
    bnnz = min(int(bnnz), nbr * nbc)
    pos = np.sort(rng.choice(nbr * nbc, size=bnnz, replace=False))
    brow, bcol = pos // nbc, pos % nbc
    indptr = np.zeros(nbr + 1, dtype=np.int64)
    np.add.at(indptr, brow + 1, 1)
    indptr = np.cumsum(indptr)
    data = rng.random((bnnz, bm, bn))
    return sp.bsr_matrix((data, bcol, indptr), shape=(nrows, ncols))
