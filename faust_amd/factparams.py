"""Parameter classes of the factorization algorithms (pyfaust.factparams
counterpart; wrapper/python/pyfaust/factparams.py).  Only the constraint
types on the hot path are provided: sp, spcol, splin, splincol, supp, and
normcol / normlin (the reference's PALM4MSA known-answer test uses normcol)."""
import os

import numpy as np

from ._lib import PROX


class StoppingCriterion:
    """factparams.py:20-31 (num_its defaults to 500; tol switches to an error criterion)."""
    DEFAULT_MAXITER = 10000
    DEFAULT_NUMITS = 500

    def __init__(self, num_its=DEFAULT_NUMITS, tol=None, maxiter=DEFAULT_MAXITER, relerr=False, relmat=None,
                 erreps=None):
        self.num_its = int(num_its)
        self.tol = tol
        self.maxiter = int(maxiter)
        self.relerr = relerr
        self.relmat = relmat
        if erreps is not None:
            raise NotImplementedError("erreps is not supported")
        if tol is not None and relerr:
            if relmat is None:
                raise ValueError("relerr=True needs relmat")
            self.tol = float(tol) * float(np.linalg.norm(relmat))

    def _c(self):
        from ._lib import fh_stop_crit
        s = fh_stop_crit()
        s.num_its = self.num_its
        s.is_criterion_error = int(self.tol is not None)
        s.error_threshold = float(self.tol) if self.tol is not None else 0.0
        s.max_num_its = self.maxiter
        return s


class ConstraintName:
    SP, SPCOL, SPLIN, SPLINCOL, SUPP = "sp", "spcol", "splin", "splincol", "supp"
    NORMCOL, NORMLIN = "normcol", "normlin"

    def __init__(self, name):
        if isinstance(name, ConstraintName):
            name = name.name
        if name not in PROX:
            raise ValueError(f"constraint {name!r} is not supported "
                             "(sp, spcol, splin, splincol, supp, normcol, normlin)")
        self.name = name


class ConstraintGeneric:
    def __init__(self, name, num_rows, num_cols, cons_value, normalized=True, pos=False):
        self._name = ConstraintName(name)
        self._num_rows = int(num_rows)
        self._num_cols = int(num_cols)
        self._cons_value = cons_value
        self.normalized = bool(normalized)
        self.pos = bool(pos)

    @property
    def name(self):
        return self._name

    @property
    def shape(self):
        return (self._num_rows, self._num_cols)

    def _c(self, keep):
        from ._lib import fh_constraint
        import ctypes as C
        c = fh_constraint()
        c.name = PROX[self._name.name]
        c.nrows, c.ncols = self._num_rows, self._num_cols
        c.normalized, c.pos = int(self.normalized), int(self.pos)
        if self._name.name == "supp":
            S = np.asfortranarray(self._cons_value, dtype=np.float64)
            keep.append(S)
            c.supp = S.ctypes.data_as(C.POINTER(C.c_double))
            c.k = 0
        elif self._name.name in ("normcol", "normlin"):
            c.k = 0
            c.s = float(self._cons_value)
        else:
            c.k = int(self._cons_value)
        return c


class ConstraintInt(ConstraintGeneric):
    """sp / spcol / splin / splincol with an integer k."""

    def __init__(self, name, num_rows, num_cols, cons_value, normalized=True, pos=False):
        super().__init__(name, num_rows, num_cols, int(cons_value), normalized, pos)


class ConstraintReal(ConstraintGeneric):
    """normcol / normlin with a real s, the 2-norm imposed on every column /
    row (factparams.py:700-745; not normalized by default)."""

    def __init__(self, name, num_rows, num_cols, cons_value, normalized=False, pos=False):
        super().__init__(name, num_rows, num_cols, float(cons_value), normalized, pos)
        if self._name.name not in ("normcol", "normlin"):
            raise ValueError("ConstraintReal is for normcol / normlin")


class ConstraintMat(ConstraintGeneric):
    """supp with a 0/1 support matrix."""

    def __init__(self, name, cons_value, normalized=True, pos=False):
        cons_value = np.asarray(cons_value, dtype=np.float64)
        super().__init__(name, cons_value.shape[0], cons_value.shape[1], cons_value, normalized, pos)


class ConstraintList:
    """ConstraintList('splin', 5, 500, 32, 'normcol', 1.0, 32, 32, ...): (name, value, nrows, ncols) groups."""

    def __init__(self, *args):
        self.clist = []
        i = 0
        while i < len(args):
            name, k, nr, nc = args[i:i + 4]
            if name in ("normcol", "normlin"):
                self.clist.append(ConstraintReal(name, nr, nc, k))
            else:
                self.clist.append(ConstraintInt(name, nr, nc, k))
            i += 4

    def __iter__(self):
        return iter(self.clist)

    def __len__(self):
        return len(self.clist)


def _constraints(cons):
    if isinstance(cons, ConstraintList):
        return cons.clist
    out = []
    for c in cons:
        out.append(c.constraint if hasattr(c, "constraint") else c)
    return out


class ParamsFact:
    """Common defaults, factparams.py:955-962.  grad_calc_opt_mode (2016
    backend only, factparams.py:952-962): DISABLED_OPT, INTERNAL_OPT or
    EXTERNAL_OPT (pyfaust's default)."""

    DISABLED_OPT = 0
    INTERNAL_OPT = 1
    EXTERNAL_OPT = 2

    def __init__(self, is_update_way_R2L=False, init_lambda=1.0, step_size=1e-16, constant_step_size=False,
                 is_verbose=False, factor_format="dynamic", packing_RL=True, no_normalization=False,
                 no_lambda=False, norm2_max_iter=100, norm2_threshold=1e-6, grad_calc_opt_mode=2, **kw):
        if grad_calc_opt_mode not in (0, 1, 2):
            raise ValueError("grad_calc_opt_mode must be DISABLED_OPT, INTERNAL_OPT or EXTERNAL_OPT")
        self.grad_calc_opt_mode = int(grad_calc_opt_mode)
        if factor_format not in ("dense", "sparse", "dynamic"):
            raise ValueError("factor_format must be either 'dense', 'sparse' or 'dynamic'")
        self.is_update_way_R2L = bool(is_update_way_R2L)
        self.init_lambda = float(init_lambda)
        self.step_size = float(step_size)
        self.constant_step_size = bool(constant_step_size)
        self.is_verbose = is_verbose
        self.factor_format = factor_format
        self.packing_RL = bool(packing_RL)
        self.no_normalization = bool(no_normalization)
        self.no_lambda = bool(no_lambda)
        self.norm2_max_iter = int(norm2_max_iter) if norm2_max_iter else 100
        self.norm2_threshold = float(norm2_threshold) if norm2_threshold else 1e-6

    def _c(self, record_traj=False, is_fact_side_left=False):
        from ._lib import fh_palm_params
        p = fh_palm_params()
        p.is_update_way_R2L = int(self.is_update_way_R2L)
        p.factor_format = {"dense": 0, "sparse": 1, "dynamic": 2}[self.factor_format]
        p.packing_RL = int(self.packing_RL)
        p.no_normalization = int(self.no_normalization)
        p.no_lambda = int(self.no_lambda)
        p.norm2_max_iter = self.norm2_max_iter
        p.norm2_threshold = self.norm2_threshold
        p.constant_step_size = int(self.constant_step_size)
        p.step_size = self.step_size
        p.is_fact_side_left = int(is_fact_side_left)
        p.record_traj = int(record_traj)
        # palm4msa2 reads USE_GRAD1 from the environment (faust_palm4msa2020.hpp:83-86)
        g1 = getattr(self, "use_grad1", None)
        if g1 is None:
            try:
                g1 = bool(int(os.environ.get("USE_GRAD1", "0")))
            except ValueError:
                g1 = False
        p.use_grad1 = int(g1)
        p.grad_calc_opt_mode = int(getattr(self, "grad_calc_opt_mode", 2))
        return p


class ParamsPalm4MSA(ParamsFact):
    """pyfaust ParamsPalm4MSA(constraints, stop_crit, init_facts=None, ...)."""

    def __init__(self, constraints, stop_crit, init_facts=None, init_lambda=1.0, step_size=1e-16,
                 constant_step_size=False, is_update_way_R2L=False, is_verbose=False, norm2_max_iter=100,
                 norm2_threshold=1e-6, factor_format="dynamic", packing_RL=True, no_normalization=False,
                 no_lambda=False, grad_calc_opt_mode=2, **kw):
        super().__init__(is_update_way_R2L, init_lambda, step_size, constant_step_size, is_verbose, factor_format,
                         packing_RL, no_normalization, no_lambda, norm2_max_iter, norm2_threshold,
                         grad_calc_opt_mode)
        self.constraints = _constraints(constraints)
        self.stop_crit = stop_crit
        self.init_facts = init_facts
        self.num_facts = len(self.constraints)


class ParamsHierarchical(ParamsFact):
    """pyfaust ParamsHierarchical(fact_constraints, res_constraints, stop_crit1, stop_crit2, ...)."""

    def __init__(self, fact_constraints, res_constraints, stop_crit1, stop_crit2, is_update_way_R2L=False,
                 init_lambda=1.0, step_size=1e-16, constant_step_size=False, is_fact_side_left=False,
                 is_verbose=False, factor_format="dynamic", packing_RL=True, no_normalization=False,
                 no_lambda=False, norm2_max_iter=100, norm2_threshold=1e-6, grad_calc_opt_mode=2, **kw):
        super().__init__(is_update_way_R2L, init_lambda, step_size, constant_step_size, is_verbose, factor_format,
                         packing_RL, no_normalization, no_lambda, norm2_max_iter, norm2_threshold,
                         grad_calc_opt_mode)
        fact_constraints = _constraints(fact_constraints)
        res_constraints = _constraints(res_constraints)
        if len(fact_constraints) != len(res_constraints):
            raise ValueError("fact_constraints and res_constraints must have the same length")
        self.fact_constraints = fact_constraints
        self.res_constraints = res_constraints
        self.stop_crits = [stop_crit1, stop_crit2]
        self.is_fact_side_left = bool(is_fact_side_left)
        self.num_facts = len(fact_constraints) + 1


class ParamsHierarchicalWHT(ParamsHierarchical):
    """Hadamard factorization parameters (factparams.py:1605-1648): n-1 levels of
    splincol k=2 factors and splincol k=d/2^(i+1) residuals, R2L, no packing."""

    def __init__(self, n, proj_name="splincol", num_its=30):
        if proj_name != "splincol":
            raise ValueError("only the splincol projector is supported")
        d = 2 ** int(n)
        stop = StoppingCriterion(num_its=num_its)
        super().__init__([ConstraintInt("splincol", d, d, 2) for _ in range(n - 1)],
                         [ConstraintInt("splincol", d, d, int(d / 2.0 ** (i + 1))) for i in range(n - 1)],
                         stop, stop, is_update_way_R2L=True, packing_RL=False)
