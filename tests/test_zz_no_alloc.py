"""fh_faust_multiply_noalloc (intermediates in the library's fixed per-GPU
scratch array) against fh_faust_multiply (the Faust's own intermediates):
both run the same kernels, so the outputs must agree bit for bit.  Each chain
kind is multiplied twice in a row through the C ABI, the second time with
fewer columns, so the second call runs over what the first left in the
array.  Products whose intermediates exceed the array are refused on the host
with FH_ERR_ARG."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import faust_amd
from faust_amd import _lib
from faust_amd.bsr import rand_bsr

pytestmark = pytest.mark.gpu


def _call(fn, F, op, X, p, Y):
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return fn(F.h, op.encode(), C.c_void_p(X.data_ptr()), p, C.c_void_p(Y.data_ptr()), st)


def _both(L, F, op, rows_in, rows_out, p, dtype, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    X = torch.randn((p, rows_in), generator=g, device="cuda", dtype=torch.float32).to(dtype)
    Y0 = torch.full((p, rows_out), float("nan"), device="cuda", dtype=dtype)
    Y1 = torch.full((p, rows_out), float("nan"), device="cuda", dtype=dtype)
    assert _call(L.fh_faust_multiply, F, op, X, p, Y0) == _lib.FH_OK, L.fh_last_error()
    rc = _call(L.fh_faust_multiply_noalloc, F, op, X, p, Y1)
    assert rc == _lib.FH_OK, L.fh_last_error()
    torch.cuda.synchronize()
    assert torch.isfinite(Y0.float()).all()
    assert torch.equal(Y0.view(torch.int16) if dtype == torch.bfloat16 else Y0,
                       Y1.view(torch.int16) if dtype == torch.bfloat16 else Y1)


def _csr(n, k, rng, dtype):
    return faust_amd.rand_csr(n, n, k, rng, dtype=dtype)


@pytest.mark.parametrize("op", ["N", "T"])
def test_noalloc_fp64_csr_chain(gpu, op):
    print("noalloc FP64 CSR 1024^2 x5, p=64 then 16", flush=True)
    rng = np.random.default_rng(1)
    F = faust_amd.Faust([_csr(1024, 8, rng, np.float64) for _ in range(5)])
    for p, seed in ((64, 1), (16, 2)):
        _both(_lib.lib(), F, op, 1024, 1024, p, torch.float64, seed)


def test_noalloc_fp64_mixed_dense_chain(gpu):
    print("noalloc FP64 dense+CSR 512x384x256, p=40 then 8", flush=True)
    rng = np.random.default_rng(2)
    facs = [rng.standard_normal((512, 384)), sp.csr_matrix(sp.random(384, 256, 0.05, random_state=3)),
            rng.standard_normal((256, 256))]
    F = faust_amd.Faust(facs)
    for p, seed in ((40, 3), (8, 4)):
        _both(_lib.lib(), F, "N", 256, 512, p, torch.float64, seed)


@pytest.mark.parametrize("op", ["N", "T"])
def test_noalloc_bsr_bf16_chain(gpu, op):
    print("noalloc BSR BF16 2048^2 x3, p=256 then 40", flush=True)
    rng = np.random.default_rng(3)
    facs = [rand_bsr(2048, 2048, 32, 4, rng, scale=1 / np.sqrt(128)) for _ in range(3)]
    F = faust_amd.Faust(facs, dtype="bf16")
    for p, seed in ((256, 5), (40, 6)):
        _both(_lib.lib(), F, op, 2048, 2048, p, torch.bfloat16, seed)


def test_noalloc_fp32_csr_chain(gpu):
    print("noalloc FP32 CSR 4096^2 x3, p=128 then 20", flush=True)
    rng = np.random.default_rng(4)
    F = faust_amd.Faust([_csr(4096, 12, rng, np.float32) for _ in range(3)], dtype=np.float32)
    for p, seed in ((128, 7), (20, 8)):
        _both(_lib.lib(), F, "N", 4096, 4096, p, torch.float32, seed)


def test_noalloc_refuses_oversized_product(gpu):
    # two 8192 x 8192 FP64 CSR factors, p = 1024: one 64 MiB intermediate
    rng = np.random.default_rng(5)
    F = faust_amd.Faust([_csr(8192, 2, rng, np.float64) for _ in range(2)])
    L = _lib.lib()
    X = torch.zeros((1024, 8192), device="cuda", dtype=torch.float64)
    Y = torch.zeros((1024, 8192), device="cuda", dtype=torch.float64)
    rc = _call(L.fh_faust_multiply_noalloc, F, "N", X, 1024, Y)
    assert rc == _lib.FH_ERR_ARG
    assert b"scratch array" in L.fh_last_error()
    # the same product with 256 columns fits
    _both(L, F, "N", 8192, 8192, 256, torch.float64, 9)
