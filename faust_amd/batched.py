"""Batched PALM4MSA on the GPU: many independent factorizations of uneven
sizes advanced in lock-step, FP32 iterations + FP64 frozen-support
refinement, one HIP-graph replay per iteration (fh_palm4msa_batched)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


class fh_batch_params(C.Structure):
    _fields_ = [("dtype", C.c_int), ("num_its", C.c_int), ("refine_its", C.c_int), ("prox", C.c_int),
                ("warm_power", C.c_int), ("norm2_max_iter", C.c_int), ("norm2_threshold", C.c_double),
                ("use_graph", C.c_int), ("autotune", C.c_int), ("check_every", C.c_int),
                ("target_err", C.c_double), ("target_errs", C.c_void_p)]


class BatchResult:
    def __init__(self, lambdas, relerr, factors, times_ms, iters_to_target):
        self.lambdas = lambdas
        self.relerr = relerr
        self.factors = factors
        # set-up, main iterations, refinement, verification, time until all targets reached target_err
        self.times_ms = times_ms
        self.iters_to_target = iters_to_target


def palm4msa_batched(targets, nfacts, k=None, supports=None, init=None, lambda_init=None, dtype="f32",
                     num_its=200, refine_its=20, warm_power=True, use_graph=True, autotune=True,
                     norm2_max_iter=100, norm2_threshold=1e-6, return_factors=True, check_every=0,
                     target_err=0.0):
    """targets: list of square FP64 arrays.  k: sp parameter per target (list or
    int) for the sp projector; supports: per target a list of nfacts 0/1 masks
    (prox supp) instead.  init: per target a list of nfacts initial factors.
    target_err: the FP64 relative error to reach (checked every check_every
    iterations), one for all targets or one per target."""
    L = lib()
    nt = len(targets)
    As = [np.asfortranarray(a, dtype=np.float64) for a in targets]
    n = (C.c_int * nt)(*[a.shape[0] for a in As])
    for a in As:
        if a.shape[0] != a.shape[1]:
            raise ValueError("batched PALM4MSA: square targets only")
    Aptr = (C.c_void_p * nt)(*[a.ctypes.data for a in As])
    keep = []
    kk = None
    sup = None
    p = fh_batch_params()
    p.dtype = _lib.FH_F32 if dtype == "f32" else _lib.FH_F64
    p.num_its, p.refine_its = int(num_its), int(refine_its)
    p.warm_power, p.use_graph, p.autotune = int(warm_power), int(use_graph), int(autotune)
    p.norm2_max_iter, p.norm2_threshold = int(norm2_max_iter), float(norm2_threshold)
    p.check_every = int(check_every)
    if np.isscalar(target_err):
        p.target_err = float(target_err)
    else:
        te = np.ascontiguousarray(target_err, dtype=np.float64)
        if te.shape != (nt,):
            raise ValueError("target_err: one value, or one per target")
        keep.append(te)
        p.target_errs = te.ctypes.data
    if supports is not None:
        p.prox = _lib.PROX["supp"]
        masks = [np.asfortranarray(m, dtype=np.float64) for s in supports for m in s]
        keep += masks
        sup = (C.c_void_p * len(masks))(*[m.ctypes.data for m in masks])
    else:
        p.prox = _lib.PROX["sp"]
        kl = [int(k)] * nt if np.isscalar(k) else [int(x) for x in k]
        kk = (C.c_int64 * nt)(*kl)
    ini = None
    if init is not None:
        facs = [np.asfortranarray(f, dtype=np.float64) for s in init for f in s]
        keep += facs
        ini = (C.c_void_p * len(facs))(*[f.ctypes.data for f in facs])
    lam = None
    if lambda_init is not None:
        lam = (C.c_double * nt)(*[float(x) for x in lambda_init])
    out = C.c_void_p()
    check(L.fh_palm4msa_batched(nt, nfacts, n, Aptr, kk, sup, ini, lam, C.byref(p), C.byref(out)))
    try:
        lams = np.zeros(nt)
        errs = np.zeros(nt)
        ms = np.zeros(5)
        iters = np.zeros(nt, dtype=np.int32)
        check(L.fh_batch_result_iters(out, iters.ctypes.data_as(C.c_void_p)))
        check(L.fh_batch_result_lambda(out, lams.ctypes.data_as(C.c_void_p)))
        check(L.fh_batch_result_relerr(out, errs.ctypes.data_as(C.c_void_p)))
        check(L.fh_batch_result_times(out, ms.ctypes.data_as(C.c_void_p)))
        facs = None
        if return_factors:
            facs = []
            for b in range(nt):
                fb = []
                for f in range(nfacts):
                    d = np.empty((As[b].shape[0], As[b].shape[0]), order="F")
                    check(L.fh_batch_result_factor(out, b, f, d.ctypes.data_as(C.c_void_p)))
                    fb.append(d)
                facs.append(fb)
    finally:
        L.fh_batch_result_free(out)
    return BatchResult(lams, errs, facs, ms, iters)
