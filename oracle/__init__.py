"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package
(faust_amd) never imports it.

Factor descriptions accepted everywhere below:
  * numpy 2-D float64 array                 -> dense factor (MatDense)
  * scipy.sparse.csr_matrix (float64)        -> sparse factor (MatSparse, CSR int32)
  * scipy.sparse.bsr_matrix (float64)        -> BSR factor (MatBSR, blocks passed
                                                column-major as in faust_MatBSR.h:120-150)
"""
import ctypes as C
import os
import subprocess

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

OR_DENSE, OR_SPARSE, OR_BSR = 0, 1, 2
CONS = {"sp": 0, "spcol": 1, "splin": 2, "splincol": 3, "supp": 4, "normcol": 5, "normlin": 6}
# projectors whose parameter is a real (the norm imposed on each column / row)
REAL_CONS = ("normcol", "normlin")


class or_fac(C.Structure):
    _fields_ = [("kind", C.c_int), ("nrows", C.c_int), ("ncols", C.c_int),
                ("dense", C.POINTER(C.c_double)),
                ("nnz", C.c_int), ("rowptr", C.POINTER(C.c_int)),
                ("colind", C.POINTER(C.c_int)), ("vals", C.POINTER(C.c_double)),
                ("bm", C.c_int), ("bn", C.c_int), ("bnnz", C.c_int),
                ("browptr", C.POINTER(C.c_int)), ("bcolinds", C.POINTER(C.c_int)),
                ("bdata", C.POINTER(C.c_double))]


class or_cons(C.Structure):
    _fields_ = [("name", C.c_int), ("k", C.c_int), ("nrows", C.c_int),
                ("ncols", C.c_int), ("normalized", C.c_int), ("pos", C.c_int),
                ("supp", C.POINTER(C.c_double)), ("s", C.c_double)]


class or_stop(C.Structure):
    _fields_ = [("num_its", C.c_int), ("is_criterion_error", C.c_int),
                ("error_threshold", C.c_double), ("max_num_its", C.c_int)]


class or_params(C.Structure):
    _fields_ = [("is_update_way_R2L", C.c_int), ("factor_format", C.c_int),
                ("packing_RL", C.c_int), ("no_normalization", C.c_int),
                ("no_lambda", C.c_int), ("norm2_max_iter", C.c_int),
                ("norm2_threshold", C.c_double), ("constant_step_size", C.c_int),
                ("step_size", C.c_double), ("is_fact_side_left", C.c_int), ("use_grad1", C.c_int),
                ("grad_calc_opt_mode", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        l = C.CDLL(_LIB_PATH)
        l.or_last_error.restype = C.c_char_p
        l.or_reduce_sum.restype = C.c_double
        l.or_reduce_sum.argtypes = [C.POINTER(C.c_double), C.c_int64]
        l.or_reduce_dot.restype = C.c_double
        l.or_reduce_dot.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64]
        l.or_result_lambda.restype = C.c_double
        _lib = l
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().or_last_error().decode())


class _FacPack:
    """Keeps numpy buffers alive while the or_fac array is in use."""

    def __init__(self, factors):
        self.keep = []
        self.arr = (or_fac * len(factors))()
        for i, f in enumerate(factors):
            d = self.arr[i]
            if sp.isspmatrix_bsr(f) or isinstance(f, sp.bsr_array):
                f = sp.bsr_matrix(f)
                f.sort_indices()
                R, Cb = f.blocksize
                data = np.ascontiguousarray(np.transpose(f.data, (0, 2, 1)), dtype=np.float64)
                bp = np.ascontiguousarray(f.indptr, dtype=np.int32)
                bc = np.ascontiguousarray(f.indices, dtype=np.int32)
                self.keep += [data, bp, bc]
                d.kind, d.nrows, d.ncols = OR_BSR, f.shape[0], f.shape[1]
                d.bm, d.bn, d.bnnz = R, Cb, len(bc)
                d.browptr, d.bcolinds, d.bdata = _ip(bp), _ip(bc), _dp(data)
            elif sp.issparse(f):
                f = sp.csr_matrix(f)
                f.sort_indices()
                rp = np.ascontiguousarray(f.indptr, dtype=np.int32)
                ci = np.ascontiguousarray(f.indices, dtype=np.int32)
                v = np.ascontiguousarray(f.data, dtype=np.float64)
                self.keep += [rp, ci, v]
                d.kind, d.nrows, d.ncols, d.nnz = OR_SPARSE, f.shape[0], f.shape[1], len(v)
                d.rowptr, d.colind, d.vals = _ip(rp), _ip(ci), _dp(v)
            else:
                a = np.asfortranarray(f, dtype=np.float64)
                self.keep.append(a)
                d.kind, d.nrows, d.ncols = OR_DENSE, a.shape[0], a.shape[1]
                d.dense = _dp(a)


def reduce_sum(x):
    x = np.ascontiguousarray(x, dtype=np.float64).ravel(order="K")
    return lib().or_reduce_sum(_dp(x), x.size)


def reduce_dot(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    y = np.ascontiguousarray(y, dtype=np.float64).ravel()
    return lib().or_reduce_dot(_dp(x), _dp(y), x.size)


def gemm(A, B, opA="N", opB="N", alpha=1.0, beta=0.0, Cin=None):
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    m = A.shape[0] if opA == "N" else A.shape[1]
    k = A.shape[1] if opA == "N" else A.shape[0]
    n = B.shape[1] if opB == "N" else B.shape[0]
    Cm = np.zeros((m, n), order="F") if Cin is None else np.asfortranarray(Cin, dtype=np.float64).copy(order="F")
    _check(lib().or_gemm(C.c_char(opA.encode()), C.c_char(opB.encode()), m, n, k, C.c_double(alpha),
                         _dp(A), A.shape[0], _dp(B), B.shape[0], C.c_double(beta), _dp(Cm), m))
    return Cm


def chain_multiply(factors, X, op="N", nthreads=0):
    """Y = op(F) @ X with F = factors[0] @ ... @ factors[-1]."""
    pk = _FacPack(factors)
    X = np.asfortranarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X.reshape(-1, 1, order="F")
    out_rows = factors[0].shape[0] if op == "N" else factors[-1].shape[1]
    Y = np.empty((out_rows, X.shape[1]), order="F")
    _check(lib().or_chain_multiply(pk.arr, len(factors), C.c_char(op.encode()), _dp(X), X.shape[1], _dp(Y),
                                   int(nthreads)))
    return Y


def get_product(factors):
    pk = _FacPack(factors)
    out = np.empty((factors[0].shape[0], factors[-1].shape[1]), order="F")
    _check(lib().or_get_product(pk.arr, len(factors), _dp(out)))
    return out


def dynprog_tables(factors):
    """(cost, split) n x n tables of dynprog_multiply for the chain."""
    pk = _FacPack(factors)
    n = len(factors)
    cost = np.zeros((n, n), dtype=np.int32)
    split = np.zeros((n, n), dtype=np.int32)
    _check(lib().or_dynprog_tables(pk.arr, n, cost.ctypes.data_as(C.c_void_p), split.ctypes.data_as(C.c_void_p)))
    return cost, split


def spectral_norm(factors, maxit=100, thr=1e-6):
    pk = _FacPack(factors)
    nrm = C.c_double()
    flag = C.c_int()
    _check(lib().or_spectral_norm(pk.arr, len(factors), maxit, C.c_double(thr), C.byref(nrm), C.byref(flag)))
    return nrm.value, flag.value


def power_x0(n):
    out = np.empty(n)
    lib().or_power_iteration_x0(n, _dp(out))
    return out


def make_cons(name, k, nrows, ncols, normalized=True, pos=False, supp=None):
    """k is the integer parameter, or the real norm s for normcol / normlin."""
    c = or_cons()
    c.name, c.nrows, c.ncols = CONS[name], nrows, ncols
    if name in REAL_CONS:
        c.k, c.s = 0, float(k)
    else:
        c.k = int(k)
    c.normalized, c.pos = int(normalized), int(pos)
    keep = None
    if supp is not None:
        keep = np.asfortranarray(supp, dtype=np.float64)
        c.supp = _dp(keep)
    return c, keep


def prox(name, M, k=0, normalized=True, pos=False, supp=None):
    M = np.asfortranarray(M, dtype=np.float64).copy(order="F")
    c, keep = make_cons(name, k, M.shape[0], M.shape[1], normalized, pos, supp)
    _check(lib().or_prox(C.byref(c), _dp(M)))
    return M


def mul_all_ends(factors, flags="N", alpha=1.0, beta=0.0, out=None):
    """multiply_order_opt_all_ends (faust_prod_opt.hpp:49-86): alpha times the
    product of op(F_i) in the reference's greedy end-pair order, plus beta * out."""
    if len(flags) != len(factors) and flags != "N":
        raise ValueError("flags: one per factor, or the single flag 'N'")
    pk = _FacPack(factors)
    dims = [(f.shape[1], f.shape[0]) if (flags[i] if len(flags) > i else flags[0]) in "TH" else f.shape
            for i, f in enumerate(factors)]
    m, n = dims[0][0], dims[-1][1]
    o = np.zeros((m, n), order="F") if out is None else np.asfortranarray(out, dtype=np.float64).copy(order="F")
    fl = flags.encode()
    _check(lib().or_mul_all_ends(pk.arr, len(factors), fl, len(fl), C.c_double(alpha), C.c_double(beta), m, n, _dp(o)))
    return o


def project_gen_kind(name, M, k=0, normalized=True, pos=False, supp=None):
    """'dense' or 'sparse': the format ConstraintGeneric::project_gen returns
    (faust_prox_gen.hpp: sparse_size > dense_size -> dense)."""
    M = np.asfortranarray(M, dtype=np.float64)
    c, keep = make_cons(name, k, M.shape[0], M.shape[1], normalized, pos, supp)
    kind = C.c_int()
    _check(lib().or_project_gen_kind(C.byref(c), _dp(M), C.byref(kind)))
    return "dense" if kind.value == OR_DENSE else "sparse"


def _params(**kw):
    p = or_params()
    p.is_update_way_R2L = int(kw.get("is_update_way_R2L", False))
    fmt = kw.get("factor_format", "dynamic")
    p.factor_format = {"dense": 0, "sparse": 1, "dynamic": 2}[fmt] if isinstance(fmt, str) else int(fmt)
    p.packing_RL = int(kw.get("packing_RL", True))
    p.no_normalization = int(kw.get("no_normalization", False))
    p.no_lambda = int(kw.get("no_lambda", False))
    p.norm2_max_iter = int(kw.get("norm2_max_iter", 100))
    p.norm2_threshold = float(kw.get("norm2_threshold", 1e-6))
    p.constant_step_size = int(kw.get("constant_step_size", False))
    p.step_size = float(kw.get("step_size", 1e-16))
    p.is_fact_side_left = int(kw.get("is_fact_side_left", False))
    p.use_grad1 = int(kw.get("use_grad1", False))
    p.grad_calc_opt_mode = int(kw.get("grad_calc_opt_mode", 2))
    return p


def _stop(num_its=500, tol=None, maxiter=10000):
    s = or_stop()
    s.num_its = int(num_its)
    s.is_criterion_error = int(tol is not None)
    s.error_threshold = float(tol) if tol is not None else 0.0
    s.max_num_its = int(maxiter)
    return s


class Result:
    def __init__(self, ptr):
        self.ptr = ptr
        L = lib()
        n = L.or_result_nfacts(ptr)
        self.lam = L.or_result_lambda(ptr)
        self.factors = []
        self.kinds = []
        for i in range(n):
            kind, nr, nc, nnz = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            L.or_result_fact_info(ptr, i, C.byref(kind), C.byref(nr), C.byref(nc), C.byref(nnz))
            d = np.empty((nr.value, nc.value), order="F")
            L.or_result_fact_dense(ptr, i, _dp(d))
            self.factors.append(d)
            self.kinds.append(kind.value)
        tl = L.or_result_traj_len(ptr)
        t = np.empty((tl, 4))
        if tl:
            L.or_result_traj(ptr, _dp(t))
        self.traj = t
        L.or_result_free(ptr)


def _cons_array(cons_list):
    arr = (or_cons * len(cons_list))()
    keep = []
    for i, c in enumerate(cons_list):
        name, k, nr, nc = c[0], c[1], c[2], c[3]
        normalized = c[4] if len(c) > 4 else True
        pos = c[5] if len(c) > 5 else False
        supp = c[6] if len(c) > 6 else None
        cc, kp = make_cons(name, k, nr, nc, normalized, pos, supp)
        arr[i] = cc
        keep.append(kp)
    return arr, keep


def palm4msa(A, cons, num_its=500, lambda_init=1.0, init_facts=None, record_traj=False, tol=None, maxiter=10000,
             **params):
    """cons: list of (name, k, nrows, ncols[, normalized[, pos[, supp]]]).
    tol: error threshold (StoppingCriterion's error branch) with maxiter."""
    A = np.asfortranarray(A, dtype=np.float64)
    ca, keep = _cons_array(cons)
    p = _params(**params)
    s = _stop(num_its, tol, maxiter)
    res = C.c_void_p()
    if init_facts:
        pk = _FacPack(init_facts)
        arr, n_init = pk.arr, len(init_facts)
    else:
        arr, n_init = None, 0
    _check(lib().or_palm4msa(_dp(A), A.shape[0], A.shape[1], ca, len(cons), C.byref(s), C.byref(p),
                             C.c_double(lambda_init), arr, n_init, int(record_traj), C.byref(res)))
    return Result(res)


def hierarchical(A, fac_cons, res_cons, num_its_local=500, num_its_global=500, record_traj=False, **params):
    A = np.asfortranarray(A, dtype=np.float64)
    fa, k1 = _cons_array(fac_cons)
    ra, k2 = _cons_array(res_cons)
    p = _params(**params)
    s1 = _stop(num_its_local)
    s2 = _stop(num_its_global)
    res = C.c_void_p()
    _check(lib().or_hierarchical(_dp(A), A.shape[0], A.shape[1], fa, ra, len(fac_cons), C.byref(s1), C.byref(s2),
                                 C.byref(p), int(record_traj), C.byref(res)))
    return Result(res)


def palm4msa2016(A, cons, num_its=500, lambda_init=1.0, init_facts=None, record_traj=False, **params):
    """The 2016 backend (faust_Palm4MSA.hpp): dense factors; trajectory rows
    (call, iteration, NaN, lambda)."""
    A = np.asfortranarray(A, dtype=np.float64)
    ca, keep = _cons_array(cons)
    p = _params(**params)
    s = _stop(num_its)
    res = C.c_void_p()
    if init_facts:
        pk = _FacPack(init_facts)
        arr, n_init = pk.arr, len(init_facts)
    else:
        arr, n_init = None, 0
    _check(lib().or_palm4msa2016(_dp(A), A.shape[0], A.shape[1], ca, len(cons), C.byref(s), C.byref(p),
                                 C.c_double(lambda_init), arr, n_init, int(record_traj), C.byref(res)))
    return Result(res)


def hierarchical2016(A, fac_cons, res_cons, num_its_local=500, num_its_global=500, init_lambda=1.0,
                     record_traj=False, **params):
    """The 2016 backend (faust_HierarchicalFact.hpp): dense factors, lambda
    not folded in."""
    A = np.asfortranarray(A, dtype=np.float64)
    fa, k1 = _cons_array(fac_cons)
    ra, k2 = _cons_array(res_cons)
    p = _params(**params)
    s1 = _stop(num_its_local)
    s2 = _stop(num_its_global)
    res = C.c_void_p()
    _check(lib().or_hierarchical2016(_dp(A), A.shape[0], A.shape[1], fa, ra, len(fac_cons), C.byref(s1),
                                     C.byref(s2), C.byref(p), C.c_double(init_lambda), int(record_traj),
                                     C.byref(res)))
    return Result(res)


def wht_params(n_log2, num_its=30):
    """ParamsHierarchicalWHT (wrapper/python/pyfaust/factparams.py:1605-1648)."""
    d = 2 ** n_log2
    fac = [("splincol", 2, d, d) for _ in range(n_log2 - 1)]
    res = [("splincol", int(d / 2.0 ** (i + 1)), d, d) for i in range(n_log2 - 1)]
    return fac, res, dict(is_update_way_R2L=True, packing_RL=False, num_its_local=num_its,
                          num_its_global=num_its)


def hadamard(n, normalized=True):
    """Sylvester Hadamard H[i,j] = (-1)^popcount(i&j), optionally / sqrt(n)."""
    i = np.arange(n)
    pc = np.vectorize(lambda v: bin(v).count("1"))(np.bitwise_and.outer(i, i))
    H = np.where(pc % 2 == 0, 1.0, -1.0)
    if normalized:
        H = H / np.sqrt(n)
    return np.asfortranarray(H)
