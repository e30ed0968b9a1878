"""palm4msa / hierarchical / butterfly on the GPU (pyfaust.fact counterpart,
wrapper/python/pyfaust/fact.py:798-879, :1020-1335 and :1642-1907; C++
entries FaustFact.hpp:580-675,740-860,905-926)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib
from .factparams import ParamsHierarchical, ParamsHierarchicalWHT, ParamsPalm4MSA
from .faust import Factor, Faust


class PalmResult:
    """Raw result: factor handles, lambda and the (call, it, rel_err, lambda) trajectory."""

    def __init__(self, ptr):
        L = lib()
        n, lam = C.c_int(), C.c_double()
        check(L.fh_palm_result_nfacts(ptr, C.byref(n)))
        check(L.fh_palm_result_lambda(ptr, C.byref(lam)))
        self.lam = lam.value
        self.factors = []
        for i in range(n.value):
            h = C.c_void_p()
            check(L.fh_palm_result_factor(ptr, i, C.byref(h)))
            self.factors.append(Factor(None, handle=h))
        tl = C.c_int()
        check(L.fh_palm_result_traj(ptr, None, C.byref(tl)))
        t = np.zeros((tl.value, 4))
        if tl.value:
            check(L.fh_palm_result_traj(ptr, t.ctypes.data_as(C.c_void_p), C.byref(tl)))
        self.traj = t
        L.fh_palm_result_free(ptr)

    def dense_factors(self):
        return [f.toarray() for f in self.factors]

    def kinds(self):
        return [f.info()[0] for f in self.factors]


def _nbytes(f):
    kind, _, r, c, nnz = f.info()
    return r * c * 8 if kind == _lib.FH_MAT_DENSE else nnz * 12 + (r + 1) * 4


def _faust_times_lambda(res):
    """TransformHelper(facts, lambda): lambda multiplies the factor of fewest
    bytes (faust_Transform.hpp:188-214)."""
    facs = res.dense_factors()
    sizes = [_nbytes(f) for f in res.factors]
    i = int(np.argmin(sizes))
    kinds = res.kinds()
    out = []
    import scipy.sparse as sp
    for j, (a, k) in enumerate(zip(facs, kinds)):
        if j == i:
            a = a * res.lam
        out.append(sp.csr_matrix(a) if k == _lib.FH_MAT_CSR else a)
    return Faust(out)


def _palm_raw(M, p, record_traj=False):
    M = np.asfortranarray(M, dtype=np.float64)
    keep = []
    cons = (_lib.fh_constraint * len(p.constraints))(*[c._c(keep) for c in p.constraints])
    sc = p.stop_crit._c()
    pc = p._c(record_traj)
    init_f = Faust(list(p.init_facts)) if p.init_facts is not None else None  # alive across the call
    out = C.c_void_p()
    check(lib().fh_palm4msa(M.ctypes.data_as(C.c_void_p), M.shape[0], M.shape[1], cons, len(p.constraints),
                            C.byref(sc), C.byref(pc), p.init_lambda, init_f.h if init_f is not None else None,
                            C.byref(out)))
    return PalmResult(out)


def _hier_raw(M, p, record_traj=False):
    M = np.asfortranarray(M, dtype=np.float64)
    keep = []
    n = len(p.fact_constraints)
    fc = (_lib.fh_constraint * n)(*[c._c(keep) for c in p.fact_constraints])
    rc = (_lib.fh_constraint * n)(*[c._c(keep) for c in p.res_constraints])
    s1, s2 = p.stop_crits[0]._c(), p.stop_crits[1]._c()
    pc = p._c(record_traj, p.is_fact_side_left)
    out = C.c_void_p()
    check(lib().fh_hierarchical(M.ctypes.data_as(C.c_void_p), M.shape[0], M.shape[1], fc, rc, n, C.byref(s1),
                                C.byref(s2), C.byref(pc), C.byref(out)))
    return PalmResult(out)


def _palm2016_raw(M, p, record_traj=False):
    M = np.asfortranarray(M, dtype=np.float64)
    keep = []
    cons = (_lib.fh_constraint * len(p.constraints))(*[c._c(keep) for c in p.constraints])
    sc = p.stop_crit._c()
    pc = p._c(record_traj)
    init_f = None  # alive across the call
    if p.init_facts is not None:
        init_f = Faust([np.asarray(f, dtype=np.float64) for f in p.init_facts])
    out = C.c_void_p()
    check(lib().fh_palm4msa2016(M.ctypes.data_as(C.c_void_p), M.shape[0], M.shape[1], cons, len(p.constraints),
                                C.byref(sc), C.byref(pc), p.init_lambda, init_f.h if init_f is not None else None,
                                C.byref(out)))
    return PalmResult(out)


def _hier2016_raw(M, p, record_traj=False):
    M = np.asfortranarray(M, dtype=np.float64)
    keep = []
    n = len(p.fact_constraints)
    fc = (_lib.fh_constraint * n)(*[c._c(keep) for c in p.fact_constraints])
    rc = (_lib.fh_constraint * n)(*[c._c(keep) for c in p.res_constraints])
    s1, s2 = p.stop_crits[0]._c(), p.stop_crits[1]._c()
    pc = p._c(record_traj, p.is_fact_side_left)
    out = C.c_void_p()
    check(lib().fh_hierarchical2016(M.ctypes.data_as(C.c_void_p), M.shape[0], M.shape[1], fc, rc, n, C.byref(s1),
                                    C.byref(s2), C.byref(pc), p.init_lambda, C.byref(out)))
    return PalmResult(out)


def _faust_2016_palm(res):
    """fact_palm4MSA_gen's Faust (FaustFact.hpp:380-410): each dense factor
    stays dense unless 2*nnz + rows + 1 < rows*cols (then sparse), and lambda
    goes to TransformHelper(facts, lambda): the factor of fewest bytes."""
    import scipy.sparse as sp
    mats = []
    for a in res.dense_factors():
        nnz = int(np.count_nonzero(a))
        mats.append(sp.csr_matrix(a) if 2 * nnz + a.shape[0] + 1 < a.shape[0] * a.shape[1] else a)
    sizes = [(m.nnz * 12 + (m.shape[0] + 1) * 4) if sp.issparse(m) else m.size * 8 for m in mats]
    i = int(np.argmin(sizes))
    mats[i] = mats[i] * res.lam
    return Faust(mats)


def _faust_2016_hier(res):
    """fact_hierarchical_gen's Faust (FaustFact.hpp:486-500): sparse factors
    (exact zeros dropped), lambda multiplied into the first one."""
    import scipy.sparse as sp
    mats = [sp.csr_matrix(a) for a in res.dense_factors()]
    mats[0] = mats[0] * res.lam
    return Faust(mats)


def palm4msa(M, p, ret_lambda=False, backend=2020, on_gpu=True, ret_raw=False):
    """Factorize M with PALM4MSA on the GPU: backend=2020 (palm4msa2,
    faust_palm4msa2020.hpp) or backend=2016 (Faust::Palm4MSA,
    faust_Palm4MSA.hpp; pyfaust's default).  ret_raw returns the factors as
    computed, lambda and the trajectory."""
    if backend not in (2016, 2020):
        raise ValueError("backend must be 2016 or 2020")
    if not isinstance(p, ParamsPalm4MSA):
        raise TypeError("p must be a ParamsPalm4MSA")
    res = (_palm2016_raw if backend == 2016 else _palm_raw)(M, p, record_traj=ret_raw)
    if ret_raw:
        return res
    F = _faust_2016_palm(res) if backend == 2016 else _faust_times_lambda(res)
    return (F, res.lam) if ret_lambda else F


def hierarchical(M, p, ret_lambda=False, backend=2020, on_gpu=True, ret_raw=False):
    """Hierarchical factorization on the GPU: backend=2020 (hierarchical,
    faust_hierarchical.hpp) or backend=2016 (Faust::HierarchicalFact,
    faust_HierarchicalFact.hpp; pyfaust's default).  p may be a
    ParamsHierarchical or the string 'hadamard'."""
    if backend not in (2016, 2020):
        raise ValueError("backend must be 2016 or 2020")
    if isinstance(p, str):
        if p.lower() != "hadamard":
            raise ValueError("only the 'hadamard' shortcut is supported")
        p = ParamsHierarchicalWHT(int(np.log2(M.shape[0])))
    if not isinstance(p, ParamsHierarchical):
        raise TypeError("p must be a ParamsHierarchical")
    res = (_hier2016_raw if backend == 2016 else _hier_raw)(M, p, record_traj=ret_raw)
    if ret_raw:
        return res
    F = _faust_2016_hier(res) if backend == 2016 else _faust_times_lambda(res)
    return (F, res.lam) if ret_lambda else F


_BFLY_DIR = {"left": _lib.FH_BUTTERFLY_LEFT, "right": _lib.FH_BUTTERFLY_RIGHT, "bbtree": _lib.FH_BUTTERFLY_BBTREE}


def _is_cuda_tensor(M):
    try:
        import torch
    except ImportError:
        return False
    return isinstance(M, torch.Tensor) and M.is_cuda


def _butterfly_one(M, type, perm, mul_perm):
    """One fh_butterfly call (the _FaustAlgoGen.pyx.in:592-630 binding)."""
    if type not in _BFLY_DIR:
        raise ValueError("dir argument must be 'right' or 'left'.")
    A = None if _is_cuda_tensor(M) else np.asfortranarray(M, dtype=np.float64)
    pp = None
    if perm is not None:
        if not isinstance(perm, (list, np.ndarray)) or len(perm) != M.shape[1]:
            raise ValueError("perm must be a valid permutation list")
        perm = np.asarray(perm)
        if perm.ndim != 1:
            raise ValueError("perm must be a vector")
        if (perm >= M.shape[1]).any() or (perm < 0).any():
            raise ValueError("perm contains an invalid index")
        perm = np.ascontiguousarray(perm, dtype=np.int32)
        pp = perm.ctypes.data_as(C.POINTER(C.c_int32))
    L = lib()
    h = C.c_void_p()
    if A is None:  # a CUDA tensor: factorized where it lies
        import torch
        T = M if M.dtype == torch.float64 else M.double()
        if T.is_contiguous():
            rm, ld = 1, T.stride(0)
        elif T.t().is_contiguous():
            rm, ld = 0, T.stride(1)
        else:
            T = T.contiguous()
            rm, ld = 1, T.stride(0)
        check(L.fh_butterfly_device(C.c_void_p(T.data_ptr()), T.shape[0], max(int(ld), 1), rm, _BFLY_DIR[type], pp,
                                    1 if mul_perm else 0, C.byref(h), C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    else:
        check(L.fh_butterfly(A.ctypes.data_as(C.POINTER(C.c_double)), A.shape[0], _BFLY_DIR[type], pp,
                             1 if mul_perm else 0, C.byref(h)))
    try:
        n = C.c_int()
        check(L.fh_faust_size(h, C.byref(n)))
        facs = []
        for i in range(n.value):
            fh = C.c_void_p()
            check(L.fh_faust_factor(h, i, C.byref(fh)))
            facs.append(Factor(None, handle=fh))
    finally:
        L.fh_faust_free(h)
    if _is_cuda_tensor(M):
        return Faust(_facs=facs)
    if M.dtype == np.float32:
        # computed in FP64, handed back in the input's precision
        import scipy.sparse as sp
        host = [f.to_host() for f in facs]
        return Faust([(x.astype(np.float32) if sp.issparse(x) else np.asarray(x, dtype=np.float32)) for x in host],
                     dtype=np.float32)
    return Faust(_facs=facs)


def butterfly(M, type="bbtree", perm=None, diag_opt=False, mul_perm=None):
    """Butterfly factorization of M on the GPU (pyfaust.fact.butterfly,
    fact.py:1642-1907; Faust::butterfly_hierarchical,
    faust_butterfly.hpp:545-860).

    M: square float64 / float32 ndarray of order n = 2^L >= 4, or a float64
    CUDA tensor (factorized in place on the GPU, fh_butterfly_device; an
    extension of pyfaust's ndarray-only API).  type: 'bbtree'
    (balanced tree of supports, CSR factors), 'right' or 'left' (one factor
    peeled off per step, dense factors).  perm: None, an index vector (P with
    P[perm[j], j] = 1; M P^T is factorized and P put back on the right),
    'bitrev', 'default_8', or a list of index vectors (the one with the
    smallest relative error wins).  mul_perm (default: not diag_opt)
    multiplies P into the last factor, else P is a factor of its own.
    diag_opt (pyfaust.opt_butterfly_faust) is not provided.
    Each block is a best rank-1 approximation computed in FP64 (float32
    inputs get FP32 factors)."""
    from .butterfly import bitrev_perm, default_8_perms, perm2indices
    if mul_perm is None:
        mul_perm = not diag_opt
    if mul_perm and diag_opt:
        raise ValueError("mul_perm and diag_opt option can not be both True.")
    if diag_opt:
        raise NotImplementedError("diag_opt (opt_butterfly_faust) is not provided")
    on_gpu = _is_cuda_tensor(M)
    if not on_gpu and not isinstance(M, np.ndarray):
        raise TypeError("butterfly() 1st argument must be a numpy ndarray.")
    if on_gpu:
        import torch
        if M.dtype != torch.float64:
            raise TypeError("butterfly(): CUDA tensors must be float64")
    elif M.dtype not in (np.float32, np.float64):
        raise TypeError("butterfly(): only real float32 / float64 matrices are supported")
    if M.ndim != 2:
        raise ValueError("butterfly(): input matrix/array invalid number of dimensions")
    m, n = M.shape
    l2_m, l2_n = np.log2([m, n])
    if l2_m != np.floor(l2_m) or l2_n != np.floor(l2_n) or m != n:
        raise ValueError("M must be square and its dimensions must be a power of two.")
    if isinstance(perm, str):
        if perm == "bitrev":
            return butterfly(M, type, perm=perm2indices(bitrev_perm(n)), diag_opt=diag_opt, mul_perm=mul_perm)
        if perm == "default_8":
            return butterfly(M, type, default_8_perms(n), diag_opt=diag_opt, mul_perm=mul_perm)
        raise ValueError("perm must be None, an index list, a list of them, 'bitrev' or 'default_8'")
    if isinstance(perm, (list, tuple)) and len(perm) and isinstance(perm[0], (list, tuple, np.ndarray)):
        best_err, best_F = np.inf, None
        Mh = M.cpu().numpy() if on_gpu else M
        nM = np.linalg.norm(Mh)
        for p in perm:
            F = butterfly(M, type, np.asarray(p), diag_opt=diag_opt, mul_perm=mul_perm)
            err = np.linalg.norm(F.toarray() - Mh) / nM
            if err < best_err:
                best_err, best_F = err, F
        return best_F
    return _butterfly_one(M, type, perm, mul_perm)
