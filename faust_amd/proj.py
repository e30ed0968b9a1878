"""Projector functors (pyfaust.proj counterpart; wrapper/python/pyfaust/proj.py:219-467).

Each functor projects a matrix on the GPU: proj(M) returns the projected
numpy array.  Supports follow std::partial_sort's tie semantics exactly.
"""
import ctypes as C

import numpy as np

from . import _lib
from .factparams import ConstraintInt, ConstraintMat, ConstraintReal


class proj_gen:
    def __call__(self, M):
        import torch
        from .faust import _stream
        M = np.asarray(M, dtype=np.float64)
        if M.shape != self.constraint.shape:
            raise ValueError("the matrix shape does not match the projector shape")
        d = torch.from_numpy(np.ascontiguousarray(M.T)).cuda()
        supp = None
        if self.constraint.name.name == "supp":
            supp = torch.from_numpy(np.ascontiguousarray(np.asarray(self.constraint._cons_value, dtype=np.float64).T)).cuda()
        name = self.constraint.name.name
        if name in ("normcol", "normlin"):
            _lib.check(_lib.lib().fh_prox_real(_lib.FH_F64, _lib.PROX[name], C.c_void_p(d.data_ptr()), M.shape[0],
                                               M.shape[1], float(self.constraint._cons_value),
                                               int(self.constraint.normalized), int(self.constraint.pos),
                                               _stream()))
        else:
            _lib.check(_lib.lib().fh_prox(_lib.FH_F64, _lib.PROX[name], C.c_void_p(d.data_ptr()), M.shape[0],
                                          M.shape[1], int(self.k), int(self.constraint.normalized),
                                          int(self.constraint.pos),
                                          C.c_void_p(supp.data_ptr()) if supp is not None else None, _stream()))
        return np.asfortranarray(d.cpu().numpy().T)


class sp(proj_gen):
    def __init__(self, shape, k, normalized=True, pos=False):
        self.k = int(k)
        self.constraint = ConstraintInt("sp", shape[0], shape[1], k, normalized, pos)


class spcol(proj_gen):
    def __init__(self, shape, k, normalized=True, pos=False):
        self.k = int(k)
        self.constraint = ConstraintInt("spcol", shape[0], shape[1], k, normalized, pos)


class splin(proj_gen):
    def __init__(self, shape, k, normalized=True, pos=False):
        self.k = int(k)
        self.constraint = ConstraintInt("splin", shape[0], shape[1], k, normalized, pos)


class splincol(proj_gen):
    def __init__(self, shape, k, normalized=True, pos=False):
        self.k = int(k)
        self.constraint = ConstraintInt("splincol", shape[0], shape[1], k, normalized, pos)


class supp(proj_gen):
    def __init__(self, S, normalized=True, pos=False):
        self.k = 0
        self.constraint = ConstraintMat("supp", S, normalized, pos)


class normcol(proj_gen):
    """Every column scaled to 2-norm s (pyfaust.proj.normcol, proj.py:516-552)."""

    def __init__(self, shape, s=1, normalized=False, pos=False):
        self.k = 0
        self.constraint = ConstraintReal("normcol", shape[0], shape[1], s, normalized, pos)


class normlin(proj_gen):
    """Every row scaled to 2-norm s (pyfaust.proj.normlin, proj.py:554-590)."""

    def __init__(self, shape, s=1, normalized=False, pos=False):
        self.k = 0
        self.constraint = ConstraintReal("normlin", shape[0], shape[1], s, normalized, pos)
