"""The Faust class: a product of sparse/dense factors held on the MI355X.

Mirrors the parts of pyfaust.Faust on the Faust x dense hot path
(wrapper/python/pyfaust.py: __init__ :214, __matmul__ :1092-1216, T/H,
toarray, norm(2), factors, numfactors, shape) on top of the C ABI of
libfaust_hip.so.  Every product runs on the GPU; there is no CPU path.
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import check, lib

try:
    import torch
except ImportError:  # torch is the device-memory plumbing of the Python layer
    torch = None

_DT = {np.dtype(np.float64): _lib.FH_F64, np.dtype(np.float32): _lib.FH_F32}


def _require_torch_cuda():
    if torch is None or not torch.cuda.is_available():
        raise RuntimeError("faust_amd needs a ROCm GPU visible to torch (no CPU fallback)")


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def colmajor_to_device(X, dtype=torch.float64 if torch else None):
    """numpy/torch (n, p) -> device tensor holding X column-major (a (p, n) row-major tensor)."""
    if torch is not None and isinstance(X, torch.Tensor):
        t = X.to(device="cuda", dtype=dtype)
        return t.t().contiguous()
    a = np.asarray(X)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    return torch.from_numpy(np.ascontiguousarray(a.T)).to(device="cuda", dtype=dtype)


class Factor:
    """One device matrix handle (MatDense / MatSparse / MatBSR on the GPU)."""

    def __init__(self, M, dtype=np.float64, handle=None):
        self.h = C.c_void_p()
        self._host = None  # exact host copy, kept for FP32 BSR factors (stored as a BF16 hi/lo pair)
        if handle is not None:
            self.h = handle
            return
        _require_torch_cuda()
        L = lib()
        dt = _lib.FH_BF16 if (isinstance(dtype, str) and dtype == "bf16") else _DT[np.dtype(dtype)]
        if sp.issparse(M) and M.format == "bsr":
            from .bsr import bsr_handle
            self.h = bsr_handle(M, dt)
            if dt == _lib.FH_F32:
                self._host = sp.bsr_matrix(M, dtype=np.float32)
        elif sp.issparse(M):
            S = sp.csr_matrix(M)
            S.sum_duplicates()
            S.sort_indices()
            rp = np.ascontiguousarray(S.indptr, dtype=np.int32)
            ci = np.ascontiguousarray(S.indices, dtype=np.int32)
            v = _host_values(S.data, dt)
            check(L.fh_mat_csr_create(dt, S.shape[0], S.shape[1], int(S.nnz), rp.ctypes.data_as(C.c_void_p),
                                      ci.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p), 0,
                                      C.byref(self.h)))
        else:
            a = np.asarray(M)
            v = _host_values(np.asfortranarray(a), dt)
            v = np.asfortranarray(v)
            check(L.fh_mat_dense_create(dt, a.shape[0], a.shape[1], v.ctypes.data_as(C.c_void_p),
                                        max(a.shape[0], 1), 0, C.byref(self.h)))

    def info(self):
        k, d, r, c, nnz = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int64()
        check(lib().fh_mat_info(self.h, C.byref(k), C.byref(d), C.byref(r), C.byref(c), C.byref(nnz)))
        return k.value, d.value, r.value, c.value, nnz.value

    @property
    def shape(self):
        return self.info()[2:4]

    def toarray(self):
        kind, dt, r, c, _ = self.info()
        # fh_mat_to_dense writes BSR factors (BF16 or FP32 values) as FP32
        tdt = torch.float32 if kind == _lib.FH_MAT_BSR else \
            {_lib.FH_F64: torch.float64, _lib.FH_F32: torch.float32, _lib.FH_BF16: torch.bfloat16}[dt]
        out = torch.empty((c, r), dtype=tdt, device="cuda")
        check(lib().fh_mat_to_dense(self.h, _ptr(out), _stream()))
        torch.cuda.current_stream().synchronize()
        return out.t().to(torch.float64).cpu().numpy()

    def to_host(self):
        """The factor as a host matrix (MatGeneric::toMatIOVar's content):
        numpy array (dense), scipy CSR (sparse) or scipy BSR, with float64
        values for FP64 factors and float32 otherwise (BF16 values exactly)."""
        if self._host is not None:
            return self._host.copy()
        kind, dt, r, c, nnz = self.info()
        L = lib()
        if kind == _lib.FH_MAT_DENSE:
            a = self.toarray()
            return np.asfortranarray(a if dt == _lib.FH_F64 else a.astype(np.float32))
        if kind == _lib.FH_MAT_CSR:
            rp = np.empty(r + 1, dtype=np.int32)
            ci = np.empty(max(nnz, 1), dtype=np.int32)
            raw = np.empty(max(nnz, 1), dtype={_lib.FH_F64: np.float64, _lib.FH_F32: np.float32,
                                               _lib.FH_BF16: np.uint16}[dt])
            check(L.fh_mat_tocpu(self.h, raw.ctypes.data_as(C.c_void_p), rp.ctypes.data_as(C.c_void_p),
                                 ci.ctypes.data_as(C.c_void_p)))
            vals = raw[:nnz]
            if dt == _lib.FH_BF16:
                from .bsr import bf16_bits_to_f32
                vals = bf16_bits_to_f32(vals)
            return sp.csr_matrix((vals, ci[:nnz], rp), shape=(r, c))
        # BSR, 32 x 32 blocks: nnz = bnnz * 1024
        bnnz = nnz // 1024
        bp = np.empty(r // 32 + 1, dtype=np.int32)
        bc = np.empty(max(bnnz, 1), dtype=np.int32)
        raw = np.empty(max(bnnz, 1) * 1024, dtype=np.uint16 if dt == _lib.FH_BF16 else np.float32)
        check(L.fh_mat_tocpu(self.h, raw.ctypes.data_as(C.c_void_p), bp.ctypes.data_as(C.c_void_p),
                             bc.ctypes.data_as(C.c_void_p)))
        if dt == _lib.FH_BF16:
            from .bsr import bf16_bits_to_f32
            raw = bf16_bits_to_f32(raw)
        blocks = raw[:bnnz * 1024].reshape(bnnz, 32, 32).transpose(0, 2, 1)  # stored column-major
        return sp.bsr_matrix((np.ascontiguousarray(blocks), bc[:bnnz], bp), shape=(r, c))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib._lib is not None:
            _lib._lib.fh_mat_free(self.h)
            self.h = C.c_void_p()


def _host_values(a, dt):
    if dt == _lib.FH_F64:
        return np.ascontiguousarray(a, dtype=np.float64) if a.ndim == 1 else np.asfortranarray(a, dtype=np.float64)
    if dt == _lib.FH_F32:
        return np.ascontiguousarray(a, dtype=np.float32) if a.ndim == 1 else np.asfortranarray(a, dtype=np.float32)
    # bf16: round-to-nearest-even from float32, stored as uint16
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return np.ascontiguousarray(r) if a.ndim == 1 else np.asfortranarray(r)


def _load_factors(filepath):
    """Host factors of a Faust .mat file and the dtype to upload them in:
    float32 when a factor holds single values, else float64.  FP64 chains take
    CSR and dense factors, so BSR factors of an FP64 file become CSR."""
    from .matio import read_factors
    mats = read_factors(filepath)
    single = any(np.asarray(m.data if sp.issparse(m) else m).dtype == np.float32 for m in mats)
    if not single:
        mats = [sp.csr_matrix(m) if sp.issparse(m) and m.format == "bsr" else m for m in mats]
    return mats, (np.float32 if single else np.float64)


class Faust:
    """F = S_1 S_2 ... S_J on the GPU (pyfaust.Faust counterpart, pyfaust.py:214).

    Faust(factors, dtype=np.float64): factors are numpy arrays (dense),
    scipy CSR matrices (sparse) or scipy BSR matrices.  Faust(filepath) /
    Faust(filepath=...) loads a Faust .mat file (pyfaust.py:102, Faust.load).
    """

    def __init__(self, factors=None, dtype=np.float64, _facs=None, _transposed=False, filepath=None, act_dtype=None):
        _require_torch_cuda()
        L = lib()
        if isinstance(factors, str):
            factors, filepath = None, factors
        if filepath is not None:
            if factors is not None or _facs is not None:
                raise ValueError("give either factors or a filepath")
            factors, dtype = _load_factors(filepath)
        if _facs is None:
            if factors is None or len(factors) == 0:
                raise ValueError("a Faust needs at least one factor")
            _facs = [f if isinstance(f, Factor) else Factor(f, dtype) for f in factors]
        self._facs = _facs
        self._transposed = _transposed
        self._dt = dtype
        kinds = [f.info()[:2] for f in self._facs]
        # activation (X, Y) dtype of the chain: FP64 chains -> float64, BSR chains
        # (BF16 or FP32 values) -> bfloat16 unless act_dtype asks float32 for FP32
        # BSR factors (FP32 blocks on FP32 MFMA, the reference's MatBSR<float>
        # precision), FP32 CSR chains -> float32
        self.h = C.c_void_p()
        check(L.fh_faust_create(C.byref(self.h)))
        for f in self._facs:
            check(L.fh_faust_push_back(self.h, f.h))
        if any(k == _lib.FH_MAT_BSR for k, _ in kinds):
            self.act_dtype = torch.bfloat16
            if act_dtype is not None and act_dtype not in ("bf16", torch.bfloat16):
                if act_dtype not in (np.float32, np.dtype(np.float32), "float32", torch.float32):
                    raise ValueError("act_dtype of a BSR chain: 'bf16' or float32")
                check(L.fh_faust_set_act_dtype(self.h, _lib.FH_F32))
                self.act_dtype = torch.float32
        elif all(d == _lib.FH_F64 for _, d in kinds):
            self.act_dtype = torch.float64
        else:
            self.act_dtype = torch.float32
        self.dtype = {torch.float64: np.dtype(np.float64), torch.float32: np.dtype(np.float32),
                      torch.bfloat16: np.dtype(np.float32)}[self.act_dtype]
        self._mul_mode = _lib.FH_MUL_AUTO

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib._lib is not None:
            _lib._lib.fh_faust_free(self.h)
            self.h = C.c_void_p()

    # --- shape / structure -------------------------------------------------
    @property
    def shape(self):
        r = self._facs[0].shape[0]
        c = self._facs[-1].shape[1]
        return (c, r) if self._transposed else (r, c)

    def numfactors(self):
        return len(self._facs)

    def __len__(self):
        return len(self._facs)

    def factors(self, i):
        """Dense copy of factor i (pyfaust Faust.factors returns the factor)."""
        idx = len(self._facs) - 1 - i if self._transposed else i
        a = self._facs[idx].toarray()
        return a.T if self._transposed else a

    def nnz_sum(self):
        return sum(f.info()[4] for f in self._facs)

    @property
    def T(self):
        G = Faust(_facs=self._facs, dtype=self._dt, _transposed=not self._transposed, act_dtype=self.act_dtype)
        G._set_mul_mode(self._mul_mode)
        return G

    def _set_mul_mode(self, mode):
        check(lib().fh_faust_set_mul_mode(self.h, int(mode)))
        self._mul_mode = int(mode)

    def _mul_candidates(self):
        """Execution forms worth timing for this chain (fh_faust_set_mul_mode)."""
        kinds = [f.info()[:2] for f in self._facs]
        if any(k == _lib.FH_MAT_BSR for k, _ in kinds):
            c = {f"BSR_KERNEL_{k}": _lib.FH_MUL_BSR_KERNEL(k) for k in (5, 1, 3, 2, 4)}
            c["FUSED"] = _lib.FH_MUL_FUSED  # the one-launch chain (square factors of one order)
            return c
        if all(d == _lib.FH_F64 for _, d in kinds):
            if all(k == _lib.FH_MAT_CSR for k, _ in kinds):
                # AUTO: the one-launch chain with entries loaded ahead when
                # every row fits one thread per row, else the default form
                return {"AUTO": _lib.FH_MUL_AUTO, "PER_FACTOR": _lib.FH_MUL_PER_FACTOR, "FUSED": _lib.FH_MUL_FUSED}
            return {"PER_FACTOR": _lib.FH_MUL_PER_FACTOR}
        return {"CSR_ROWS1": _lib.FH_MUL_CSR_ROWS1, "CSR_ROWS2": _lib.FH_MUL_CSR_ROWS2}

    def optimize_time(self, transp=False, inplace=False, nsamples=1, mat=None):
        """pyfaust Faust.optimize_time (pyfaust.py:3184-3240;
        TransformHelper::optimize_multiply, faust_TransformHelper.hpp:436-512):
        time each execution form of F @ mat on the GPU (nsamples products
        each, after one untimed product, HIP events on the launch stream)
        and keep the quickest.  mat defaults to a random matrix with
        min(ncols, 512) columns.  Returns F itself (inplace) or a new Faust
        on the same factors; `mul_mode_name` and `mul_mode_times_ms` record
        the choice and the measurements."""
        G = self.T if transp else self
        if mat is None:
            rng = np.random.default_rng(0)
            mat = rng.standard_normal((G.shape[1], min(G.shape[1], 512)))
        Xd = mat if (torch is not None and isinstance(mat, torch.Tensor) and mat.is_cuda) else \
            colmajor_to_device(mat, G.act_dtype)
        out = torch.empty((Xd.shape[0], G.shape[0]), dtype=Xd.dtype, device="cuda")
        times = {}
        saved = G._mul_mode
        try:
            for name, mode in self._mul_candidates().items():
                G._set_mul_mode(mode)
                G.matmul_device(Xd, out)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(max(1, int(nsamples))):
                    G.matmul_device(Xd, out)
                e1.record()
                torch.cuda.synchronize()
                times[name] = e0.elapsed_time(e1) / max(1, int(nsamples))
        finally:
            G._set_mul_mode(saved)
        best = min(times, key=times.get)
        target = self if inplace else Faust(_facs=self._facs, dtype=self._dt, _transposed=self._transposed,
                                            act_dtype=self.act_dtype)
        target._set_mul_mode(self._mul_candidates()[best])
        target.mul_mode_name = best
        target.mul_mode_times_ms = times
        return target

    @property
    def H(self):
        return self.T  # real scalars

    def transpose(self):
        return self.T

    # --- products ----------------------------------------------------------
    def _op(self):
        return b"T" if self._transposed else b"N"

    def matmul_device(self, Xd, out=None):
        """Y = F @ X on device tensors in column-major form.

        Xd: a (p, n) contiguous CUDA tensor = X (n x p) column-major.
        Returns a (p, m) CUDA tensor = Y column-major.
        """
        m = self.shape[0]
        p = Xd.shape[0]
        if Xd.shape[1] != self.shape[1]:
            raise ValueError("dimensions of the Faust and the matrix are not consistent")
        if out is None:
            out = torch.empty((p, m), dtype=Xd.dtype, device="cuda")
        check(lib().fh_faust_multiply(self.h, self._op(), _ptr(Xd), p, _ptr(out), _stream()))
        return out

    def __matmul__(self, A):
        """pyfaust Faust.__matmul__ for a dense ndarray / CUDA tensor (pyfaust.py:1177-1203)."""
        if isinstance(A, Faust):
            raise NotImplementedError("Faust @ Faust is outside the hot path")
        is_torch = torch is not None and isinstance(A, torch.Tensor)
        if not is_torch and sp.issparse(A):
            A = A.toarray()
        vec = (not is_torch and np.ndim(A) == 1) or (is_torch and A.dim() == 1)
        Xd = colmajor_to_device(A, self.act_dtype)
        Y = self.matmul_device(Xd)
        if is_torch:
            return Y.t()
        torch.cuda.current_stream().synchronize()
        y = Y.float().cpu().numpy().T if self.act_dtype == torch.bfloat16 else Y.cpu().numpy().T
        return y.reshape(-1) if vec else np.asfortranarray(y)

    def dot(self, A):
        return self @ A

    # --- files ---------------------------------------------------------------
    def save(self, filepath, format="Matlab"):
        """Write the Faust to a MATLAB v5 .mat file (pyfaust Faust.save,
        pyfaust.py:2564-2599; Transform::save_mat_file,
        faust_Transform.hpp:434-476).  A transposed Faust is written as its
        factors transposed, in reverse order, as the reference does."""
        if format != "Matlab":
            raise ValueError("Only Matlab or Matlab_core format is supported.")
        from .matio import write_factors
        mats = [f.to_host() for f in self._facs]
        if self._transposed:
            mats = [m.T if not sp.issparse(m) else (sp.bsr_matrix(m.T) if m.format == "bsr" else sp.csr_matrix(m.T))
                    for m in reversed(mats)]
        write_factors(filepath, mats)

    @staticmethod
    def load(filepath):
        """Faust from a .mat file written by Faust.save (pyfaust.py:2601-2639)."""
        return Faust(filepath=filepath)

    def toarray(self):
        """Dense product (Transform::get_product).  FP32 / BF16 chains: the
        product applied to the identity, in the chain's own precision."""
        m, n = self._facs[0].shape[0], self._facs[-1].shape[1]
        if self.act_dtype != torch.float64:
            return np.asfortranarray(np.asarray(self @ np.eye(self.shape[1]), dtype=np.float64))
        out = torch.empty((n, m), dtype=torch.float64, device="cuda")
        check(lib().fh_faust_get_product(self.h, _ptr(out), _stream()))
        a = out.t().cpu().numpy()
        return np.asfortranarray(a.T if self._transposed else a)

    def norm(self, ord="fro", **kwargs):
        """norm(2): power iteration on the chain (TransformHelper::spectralNorm);
        'fro': Frobenius norm of the dense product."""
        if ord == 2:
            if self.act_dtype != torch.float64:
                # FP32 / BF16 chains: the FP64 power iteration on an FP64 copy of the factors
                mats = [f.to_host() for f in self._facs]
                mats = [sp.csr_matrix(m, dtype=np.float64) if sp.issparse(m) else np.asarray(m, dtype=np.float64)
                        for m in mats]
                return Faust(mats).norm(2, **kwargs)
            nrm, flag = C.c_double(), C.c_int()
            check(lib().fh_faust_spectral_norm(self.h, int(kwargs.get("max_num_its", 100)),
                                                float(kwargs.get("threshold", 1e-6)), C.byref(nrm), C.byref(flag),
                                                _stream()))
            return nrm.value
        if ord == "fro":
            return float(np.linalg.norm(self.toarray()))
        raise ValueError("norm: only 2 and 'fro' are supported")

    def __repr__(self):
        return f"Faust size {self.shape[0]}x{self.shape[1]}, {self.numfactors()} factors, on MI355X"
