"""The gpu_mod function tables (include/faust_gm.h) driven the way the
reference's GPU2 classes drive them, through ctypes function pointers:

* MatSparse<GPU2>::multiply(MatDense&) (faust_MatSparse_gpu.cpp.in:462-472):
  out = spm_funcs->mul_gpu_dsm_ext(S, M, NULL, opS, OP_NOTRANSP), then
  dsm_funcs->free(M) and M = out;
* MatSparse<GPU2>::spgemm (:475-506): spm_funcs->gemm(A, B, C, &alpha, &beta, opA, opB);
* MatBSR<GPU2>::multiply / bsrgemm (faust_MatBSR_gpu.cpp.in:109-152):
  bsr_funcs->mul_gpu_dsm_ext(B, M, NULL, op, OP_NOTRANSP, NULL, NULL) and
  (A, B, C, opA, opB, &alpha, &beta);
* MatDense<GPU2>::gemm / prox_sp / prox_spcol / prox_splin / spectralNorm /
  norm (faust_MatDense_gpu.cpp.in:518-551, :746-765, :320-324, :156).
FP64 results equal the oracle bit for bit; FP32 CSR within 1e-5 relative;
FP32 BSR (BF16 activations) within 1e-2."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from faust_amd import _lib
from faust_amd.bsr import rand_bsr

pytestmark = pytest.mark.gpu

VP, I32 = C.c_void_p, C.c_int32
N, T_, H_ = 0, 1, 2


def _tables(T):
    P = C.POINTER(T)

    class Dense(C.Structure):
        _fields_ = [("create", C.CFUNCTYPE(VP, I32, I32)), ("togpu", C.CFUNCTYPE(VP, I32, I32, P)),
                    ("tocpu", C.CFUNCTYPE(None, VP, P)), ("free", C.CFUNCTYPE(None, VP)),
                    ("info", C.CFUNCTYPE(None, VP, C.POINTER(I32), C.POINTER(I32))),
                    ("gemm", C.CFUNCTYPE(None, VP, VP, VP, P, P, C.c_int, C.c_int)),
                    ("mul_gpu_dsm_ext", C.CFUNCTYPE(VP, VP, VP, VP, C.c_int, C.c_int)),
                    ("prox_sp", C.CFUNCTYPE(None, VP, I32, C.c_bool, C.c_bool)),
                    ("prox_spcol", C.CFUNCTYPE(None, VP, I32, C.c_bool, C.c_bool)),
                    ("prox_splin", C.CFUNCTYPE(None, VP, I32, C.c_bool, C.c_bool)),
                    ("prox_splincol", C.CFUNCTYPE(None, VP, I32, C.c_bool, C.c_bool)),
                    ("norm_spectral", C.CFUNCTYPE(T, VP, C.c_float, I32)), ("norm_frob", C.CFUNCTYPE(T, VP))]

    class Sparse(C.Structure):
        _fields_ = [("togpu", C.CFUNCTYPE(VP, I32, I32, I32, C.POINTER(I32), C.POINTER(I32), P)),
                    ("tocpu", C.CFUNCTYPE(None, VP, C.POINTER(I32), C.POINTER(I32), P)),
                    ("free", C.CFUNCTYPE(None, VP)),
                    ("info", C.CFUNCTYPE(None, VP, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32))),
                    ("mul_gpu_dsm_ext", C.CFUNCTYPE(VP, VP, VP, VP, C.c_int, C.c_int)),
                    ("gemm", C.CFUNCTYPE(None, VP, VP, VP, P, P, C.c_int, C.c_int))]

    class Bsr(C.Structure):
        _fields_ = [("togpu", C.CFUNCTYPE(VP, I32, I32, I32, I32, P, C.POINTER(I32), C.POINTER(I32), I32)),
                    ("free", C.CFUNCTYPE(None, VP)),
                    ("info", C.CFUNCTYPE(None, VP, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32),
                                         C.POINTER(I32), C.POINTER(I32))),
                    ("mul_gpu_dsm_ext", C.CFUNCTYPE(VP, VP, VP, VP, C.c_int, C.c_int, P, P))]

    L = _lib.lib()
    sfx = "double" if T is C.c_double else "float"
    d, s, b = Dense(), Sparse(), Bsr()
    getattr(L, f"fh_gm_load_dsm_funcs_{sfx}")(C.byref(d))
    getattr(L, f"fh_gm_load_spm_funcs_{sfx}")(C.byref(s))
    getattr(L, f"fh_gm_load_bsr_funcs_{sfx}")(C.byref(b))
    return d, s, b


def _np(T):
    return np.float64 if T is C.c_double else np.float32


def _ptr(a, T):
    return a.ctypes.data_as(C.POINTER(T))


def _dense_to(dsm, A, T):
    A = np.asfortranarray(A, dtype=_np(T))
    h = dsm.togpu(A.shape[0], A.shape[1], _ptr(A, T))
    assert h, _lib.lib().fh_last_error()
    return h


def _dense_from(dsm, h, T):
    m, n = I32(), I32()
    dsm.info(h, C.byref(m), C.byref(n))
    out = np.empty((m.value, n.value), dtype=_np(T), order="F")
    dsm.tocpu(h, _ptr(out, T))
    return out


def _sparse_to(spm, S, T):
    S = sp.csr_matrix(S)
    S.sort_indices()
    rp, ci = S.indptr.astype(np.int32), S.indices.astype(np.int32)
    v = S.data.astype(_np(T))
    h = spm.togpu(S.shape[0], S.shape[1], S.nnz, _ptr(rp, I32), _ptr(ci, I32), _ptr(v, T))
    assert h, _lib.lib().fh_last_error()
    return h


def _bsr_to(bsr, B, T):
    B = sp.bsr_matrix(B)
    B.sort_indices()
    bm, bn = B.blocksize
    data = np.ascontiguousarray(np.transpose(B.data, (0, 2, 1))).astype(_np(T)).ravel()  # column-major blocks
    rp, ci = B.indptr.astype(np.int32), B.indices.astype(np.int32)
    h = bsr.togpu(B.shape[0], B.shape[1], bm, bn, _ptr(data, T), _ptr(rp, I32), _ptr(ci, I32), B.indices.shape[0])
    assert h, _lib.lib().fh_last_error()
    return h


@pytest.mark.parametrize("op", [N, T_, H_])
def test_sparse_multiply_sequence(gpu, op):
    dsm, spm, _ = _tables(C.c_double)
    rng = np.random.default_rng(1)
    S = sp.random(300, 200, density=0.05, random_state=2, format="csr")
    X = rng.standard_normal((300 if op else 200, 17))
    hS, hM = _sparse_to(spm, S, C.c_double), _dense_to(dsm, X, C.c_double)
    # faust_MatSparse_gpu.cpp.in:462-472: new output, the input freed, the handle replaced
    out = spm.mul_gpu_dsm_ext(hS, hM, None, op, N)
    assert out and out != hM
    dsm.free(hM)
    hM = out
    got = _dense_from(dsm, hM, C.c_double)
    want = oracle.chain_multiply([S], X, "N" if op == N else "T")
    assert np.array_equal(got, want)
    # into an existing output of another shape (resized, beta = 0)
    hO = dsm.create(3, 3)
    assert spm.mul_gpu_dsm_ext(hS, _dense_to(dsm, X, C.c_double), hO, op, N) == hO
    assert np.array_equal(_dense_from(dsm, hO, C.c_double), want)
    for h in (hM, hO):
        dsm.free(h)
    spm.free(hS)
    assert _lib.lib().fh_gm_take_error() == _lib.FH_OK


def test_spgemm_alpha_beta_and_opB(gpu):
    dsm, spm, _ = _tables(C.c_double)
    rng = np.random.default_rng(3)
    A = sp.random(120, 90, density=0.1, random_state=4, format="csr")
    B = rng.standard_normal((40, 120))  # op(B) = B^T
    C0 = rng.standard_normal((90, 40))
    hA, hB, hC = _sparse_to(spm, A, C.c_double), _dense_to(dsm, B, C.c_double), _dense_to(dsm, C0, C.c_double)
    al, be = C.c_double(0.5), C.c_double(-2.0)
    spm.gemm(hA, hB, hC, C.byref(al), C.byref(be), T_, T_)
    got = _dense_from(dsm, hC, C.c_double)
    want = oracle.gemm(A.toarray(), np.ascontiguousarray(B.T), "T", "N", 0.5, -2.0, C0)
    assert np.array_equal(got, want)
    # rows/columns
    r, c, z = I32(), I32(), I32()
    spm.info(hA, C.byref(r), C.byref(c), C.byref(z))
    assert (r.value, c.value, z.value) == (120, 90, A.nnz)
    rp, ci, v = np.empty(121, np.int32), np.empty(A.nnz, np.int32), np.empty(A.nnz)
    spm.tocpu(hA, _ptr(rp, I32), _ptr(ci, I32), _ptr(v, C.c_double))
    assert np.array_equal(rp, A.indptr) and np.array_equal(ci, A.indices) and np.array_equal(v, A.data)
    for h in (hB, hC):
        dsm.free(h)
    spm.free(hA)


@pytest.mark.parametrize("op", [N, T_])
def test_bsr_multiply_and_bsrgemm(gpu, op):
    dsm, _, bsr = _tables(C.c_double)
    rng = np.random.default_rng(5)
    Bm = sp.bsr_matrix(rand_bsr(128, 96, 32, 2, rng))
    Bm2 = sp.bsr_matrix(sp.random(12, 8, density=0.5, random_state=6).toarray(), blocksize=(4, 2))
    for Bx in (Bm, Bm2):
        X = rng.standard_normal((Bx.shape[0] if op else Bx.shape[1], 9))
        hB, hM = _bsr_to(bsr, Bx, C.c_double), _dense_to(dsm, X, C.c_double)
        # faust_MatBSR_gpu.cpp.in:109-117: alpha / beta NULL
        out = bsr.mul_gpu_dsm_ext(hB, hM, None, op, N, None, None)
        dsm.free(hM)
        want = oracle.chain_multiply([Bx], X, "N" if op == N else "T")
        assert np.array_equal(_dense_from(dsm, out, C.c_double), want)
        # bsrgemm (:123-140): C given, alpha and beta
        C0 = rng.standard_normal(want.shape)
        hC = _dense_to(dsm, C0, C.c_double)
        al, be = C.c_double(1.5), C.c_double(0.25)
        assert bsr.mul_gpu_dsm_ext(hB, _dense_to(dsm, X, C.c_double), hC, op, N, C.byref(al), C.byref(be)) == hC
        ref = oracle.gemm(Bx.toarray(), X, "N" if op == N else "T", "N", 1.5, 0.25, C0)
        assert np.array_equal(_dense_from(dsm, hC, C.c_double), ref)
        for h in (out, hC):
            dsm.free(h)
        bsr.free(hB)


def test_dense_gemm_prox_norms(gpu):
    dsm, _, _ = _tables(C.c_double)
    rng = np.random.default_rng(7)
    A, B, C0 = rng.standard_normal((30, 40)), rng.standard_normal((50, 30)), rng.standard_normal((40, 50))
    hA, hB, hC = (_dense_to(dsm, M, C.c_double) for M in (A, B, C0))
    al, be = C.c_double(2.0), C.c_double(0.5)
    dsm.gemm(hA, hB, hC, C.byref(al), C.byref(be), T_, T_)
    assert np.array_equal(_dense_from(dsm, hC, C.c_double),
                          oracle.gemm(A, np.ascontiguousarray(B.T), "T", "N", 2.0, 0.5, C0))
    prod = dsm.mul_gpu_dsm_ext(hA, hA, None, T_, N)
    assert np.array_equal(_dense_from(dsm, prod, C.c_double), oracle.gemm(A, A, "T", "N"))
    for name, k in (("sp", 100), ("spcol", 4), ("splin", 5), ("splincol", 3)):
        h = _dense_to(dsm, A, C.c_double)
        getattr(dsm, "prox_" + name)(h, k, True, False)
        assert np.array_equal(_dense_from(dsm, h, C.c_double), oracle.prox(name, A, k))
        dsm.free(h)
    thr = float(np.float32(1e-6))
    assert dsm.norm_spectral(hA, thr, 100) == oracle.spectral_norm([A], 100, thr)[0]
    assert abs(dsm.norm_frob(hA) - np.linalg.norm(A)) < 1e-13 * np.linalg.norm(A)
    for h in (hA, hB, hC, prod):
        dsm.free(h)


def test_float_tables_products(gpu):
    dsm, spm, bsr = _tables(C.c_float)
    rng = np.random.default_rng(9)
    S = sp.random(256, 256, density=0.05, random_state=10, format="csr", dtype=np.float32)
    X = rng.standard_normal((256, 33)).astype(np.float32)
    hS, hM = _sparse_to(spm, S, C.c_float), _dense_to(dsm, X, C.c_float)
    out = spm.mul_gpu_dsm_ext(hS, hM, None, N, N)
    want = oracle.chain_multiply([S.astype(np.float64)], X.astype(np.float64), "N")
    got = _dense_from(dsm, out, C.c_float)
    assert np.linalg.norm(got - want) <= 1e-5 * np.linalg.norm(want)
    Bm = rand_bsr(256, 256, 32, 3, rng, scale=1 / np.sqrt(96)).astype(np.float32)
    hB = _bsr_to(bsr, Bm, C.c_float)
    outb = bsr.mul_gpu_dsm_ext(hB, hM, None, N, N, None, None)
    wantb = oracle.chain_multiply([sp.bsr_matrix(Bm).astype(np.float64)], X.astype(np.float64), "N")
    gotb = _dense_from(dsm, outb, C.c_float)
    assert np.linalg.norm(gotb - wantb) <= 1e-2 * np.linalg.norm(wantb)
    # FP64-only roles report FH_ERR_UNSUPPORTED
    L = _lib.lib()
    L.fh_gm_take_error()
    assert dsm.norm_frob(hM) == 0.0
    assert L.fh_gm_take_error() == _lib.FH_ERR_UNSUPPORTED
    for h in (hM, out, outb):
        dsm.free(h)
    spm.free(hS)
    bsr.free(hB)


def test_errors_and_general_purpose(gpu):
    dsm, spm, _ = _tables(C.c_double)
    L = _lib.lib()
    L.fh_gm_take_error()
    S = sp.random(10, 20, density=0.3, random_state=1, format="csr")
    hS, hM = _sparse_to(spm, S, C.c_double), _dense_to(dsm, np.ones((7, 3)), C.c_double)
    assert not spm.mul_gpu_dsm_ext(hS, hM, None, N, N)  # 20 != 7
    assert L.fh_gm_take_error() == _lib.FH_ERR_ARG
    assert b"dimensions" in L.fh_last_error()
    assert L.fh_gm_take_error() == _lib.FH_OK
    assert not spm.mul_gpu_dsm_ext(hS, hM, None, 7, N)  # not a gm_Op
    assert L.fh_gm_take_error() == _lib.FH_ERR_ARG

    class GP(C.Structure):
        _fields_ = [("dev_count", C.CFUNCTYPE(I32)), ("set_dev", C.CFUNCTYPE(None, I32)),
                    ("cur_dev", C.CFUNCTYPE(I32))]
    gp = GP()
    L.fh_gm_load_gp_funcs(C.byref(gp))
    assert gp.dev_count() >= 1
    gp.set_dev(0)
    assert gp.cur_dev() == 0
    dsm.free(hM)
    spm.free(hS)
