"""Faust factors <-> MATLAB v5 .mat files, in the reference's layout.

The file holds one variable, ``faust_factors``: a 1 x J cell array whose
cells are the factors in product order (faust_Transform.hpp:434-484,
save_mat_file / read_from_mat_file):
- a sparse factor is a MATLAB sparse matrix (double);
- a dense factor is a double (or single) array;
- a BSR factor is itself a 1 x 5 cell (faust_MatBSR.hpp:387-480,
  toMatIOVar): the string 'bsr', int32 [nrows, ncols, bnnz], int32 block
  column indices (1 x bnnz), int32 block row pointers (1 x nbr+1), and the
  blocks' values (1 x bnnz*bm*bn), each block column-major, one after the
  other.

Reading follows pyfaust's Faust.load (pyfaust.py:2601-2639).  Only host
arrays are handled here; faust.Faust.save / Faust.load move them to and from
the GPU.
"""
import numpy as np
import scipy.sparse as sp
from scipy.io import loadmat, savemat


def _bsr_cell(B):
    B = sp.bsr_matrix(B)
    B.sort_indices()
    bm, bn = B.blocksize
    bnnz = B.indices.shape[0]
    cell = np.empty((1, 5), dtype=object)
    cell[0, 0] = "bsr"
    cell[0, 1] = np.array([[B.shape[0], B.shape[1], bnnz]], dtype=np.int32)
    cell[0, 2] = np.asarray(B.indices, dtype=np.int32).reshape(1, -1)
    cell[0, 3] = np.asarray(B.indptr, dtype=np.int32).reshape(1, -1)
    data = B.data if B.data.dtype in (np.float32, np.float64) else B.data.astype(np.float64)
    # each (bm x bn) block column-major: element (i, j) of block b at b*bm*bn + j*bm + i
    cell[0, 4] = np.ascontiguousarray(np.transpose(data, (0, 2, 1))).reshape(1, -1)
    return cell


def _cell_of(M):
    if sp.issparse(M) and M.format == "bsr":
        return _bsr_cell(M)
    if sp.issparse(M):
        return sp.csc_matrix(M, dtype=np.float64)
    a = np.asarray(M)
    if a.ndim != 2:
        raise ValueError("a Faust factor must be a 2-D matrix")
    return a if a.dtype in (np.float32, np.float64) else a.astype(np.float64)


def write_factors(filepath, factors):
    """Write host factors (numpy arrays, scipy CSR/CSC/BSR matrices) as a
    Faust .mat file."""
    if len(factors) == 0:
        raise ValueError("can't save an empty Faust")
    cell = np.empty((1, len(factors)), dtype=object)
    for i, M in enumerate(factors):
        cell[0, i] = _cell_of(M)
    savemat(filepath, {"faust_factors": cell}, format="5", do_compression=False, oned_as="row")


def _is_bsr_cell(c):
    return (isinstance(c, np.ndarray) and c.dtype == object and c.size == 5 and isinstance(c.flat[0], np.ndarray)
            and c.flat[0].dtype.kind == "U" and c.flat[0].size == 1 and c.flat[0].flat[0] == "bsr")


def _bsr_of(c):
    nrows, ncols, bnnz = (int(x) for x in np.asarray(c.flat[1]).ravel()[:3])
    bcolinds = np.asarray(c.flat[2]).ravel().astype(np.int32)
    browptr = np.asarray(c.flat[3]).ravel().astype(np.int32)
    bdata = np.asarray(c.flat[4]).ravel()
    bnrows = nrows // (browptr.shape[0] - 1)
    bncols = bdata.shape[0] // bnrows // bnnz if bnnz else ncols
    blocks = bdata.reshape(bnnz, bncols, bnrows).transpose(0, 2, 1)  # back from column-major
    return sp.bsr_matrix((np.ascontiguousarray(blocks), bcolinds, browptr), shape=(nrows, ncols))


def read_factors(filepath):
    """Host factors of a Faust .mat file, in product order: scipy CSR for
    sparse cells, scipy BSR for BSR cells, numpy arrays for dense ones."""
    contents = loadmat(filepath)
    if "faust_factors" not in contents:
        raise ValueError(f"{filepath}: no faust_factors variable")
    cells = contents["faust_factors"]
    if cells.dtype != object or cells.ndim != 2:
        raise ValueError(f"{filepath}: faust_factors is not a cell array")
    out = []
    for c in cells.reshape(-1):
        if sp.issparse(c):
            out.append(sp.csr_matrix(c))
        elif _is_bsr_cell(c):
            out.append(_bsr_of(c))
        else:
            out.append(np.asfortranarray(np.asarray(c)))
    return out
