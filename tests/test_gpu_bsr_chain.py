"""The one-launch BSR chain (k_bsr_chain.hip) against the factor-by-factor
product (k_bsr_shared per factor): the same accumulation order and epilogue,
so the same bits, for N and T, column counts that leave some XCD teams
without columns, and FP32 (hi/lo) blocks.  Every chain call is checked
(FH_BSR_CHAIN_CHECK=1: a bad report from the kernel raises).  A tiny spin
limit forces the bounded waits to run out, from the host: the call reports
it and a later call with the normal limit is correct again."""
import numpy as np
import pytest
import torch

import faust_amd
from faust_amd import _lib
from faust_amd.bsr import rand_bsr

pytestmark = pytest.mark.gpu


def _faust(n, J, seed, dtype="bf16", bpr=4):
    rng = np.random.default_rng(seed)
    return faust_amd.Faust([rand_bsr(n, n, 32, bpr, rng, scale=1 / np.sqrt(32 * bpr)) for _ in range(J)],
                           dtype=dtype)


def _x(n, p, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.randn((p, n), generator=g, device="cuda", dtype=torch.float32).to(torch.bfloat16)


def _run(F, X, mode, monkeypatch, op="N", reps=1):
    monkeypatch.setenv("FH_BSR_CHAIN", mode)
    G = F.T if op == "T" else F
    Y = None
    for _ in range(reps):
        Y = G.matmul_device(X)
    torch.cuda.synchronize()
    L = _lib.lib()
    import ctypes as C
    launches = C.c_int()
    _lib.check(L.fh_faust_last_launches(G.h, C.byref(launches)))
    return Y.view(torch.int16).clone(), launches.value


@pytest.mark.parametrize("op", ["N", "T"])
def test_chain_cfg3a_shape_bitexact(gpu, monkeypatch, op):
    # 8192^2 x 6 factors, 32x32 blocks, 4 per block row, X 8192 x 512
    sync = "1"
    print(f"BSR chain 8192^2 x6 p=512 op={op}", flush=True)
    monkeypatch.setenv("FH_BSR_CHAIN_CHECK", "1")
    F = _faust(8192, 6, 11)
    X = _x(8192, 512, 1)
    ref, l0 = _run(F, X, "0", monkeypatch, op)
    got, l1 = _run(F, X, sync, monkeypatch, op, reps=3)
    assert l0 == 6 and l1 == 1
    assert torch.equal(ref, got)


@pytest.mark.parametrize("n,J,p", [(2048, 2, 96), (2048, 3, 40), (1024, 5, 8), (4096, 4, 1000), (512, 16, 300)])
def test_chain_shapes_bitexact(gpu, monkeypatch, n, J, p):
    sync = "1"
    print(f"BSR chain {n}^2 x{J} p={p}", flush=True)
    monkeypatch.setenv("FH_BSR_CHAIN_CHECK", "1")
    F = _faust(n, J, n + J)
    X = _x(n, p, 2)
    ref, _ = _run(F, X, "0", monkeypatch)
    got, l1 = _run(F, X, sync, monkeypatch, reps=2)
    assert l1 == 1
    assert torch.equal(ref, got)


def test_chain_fp32_blocks_bitexact(gpu, monkeypatch):
    print("BSR chain FP32 hi/lo blocks 4096^2 x4 p=256", flush=True)
    monkeypatch.setenv("FH_BSR_CHAIN_CHECK", "1")
    F = _faust(4096, 4, 5, dtype=np.float32)
    X = _x(4096, 256, 3)
    ref, _ = _run(F, X, "0", monkeypatch)
    got, l1 = _run(F, X, "1", monkeypatch, reps=2)
    assert l1 == 1
    assert torch.equal(ref, got)


def test_chain_wait_limit_reports_and_recovers(gpu, monkeypatch):
    # forced from the host: every wait asks for one item more than exists
    # (test hook) and gives up after 64 polls.  The kernel still ends,
    # reports code 1, and the next call without the hook is correct again.
    print("BSR chain forced wait limit", flush=True)
    monkeypatch.setenv("FH_BSR_CHAIN_CHECK", "1")
    F = _faust(4096, 6, 7)
    X = _x(4096, 512, 4)
    ref, _ = _run(F, X, "0", monkeypatch)
    monkeypatch.setenv("FH_BSR_CHAIN_SPIN", "64")
    monkeypatch.setenv("FH_BSR_CHAIN_TEST_UNREACHABLE", "1")
    monkeypatch.setenv("FH_BSR_CHAIN", "1")
    with pytest.raises(_lib.FaustHipError, match="reported code 1"):
        F.matmul_device(X)
        torch.cuda.synchronize()
    monkeypatch.delenv("FH_BSR_CHAIN_TEST_UNREACHABLE")
    monkeypatch.delenv("FH_BSR_CHAIN_SPIN")
    got, l1 = _run(F, X, "1", monkeypatch, reps=2)
    assert l1 == 1
    assert torch.equal(ref, got)


def test_chain_no_alloc_entry_point(gpu, monkeypatch):
    # the allocation-free entry point keeps the chain's control block in its
    # scratch array: same bits as the factor-by-factor product, twice in a row
    import ctypes as C
    print("BSR chain through fh_faust_multiply_noalloc", flush=True)
    monkeypatch.setenv("FH_BSR_CHAIN_CHECK", "1")
    F = _faust(8192, 6, 12)
    L = _lib.lib()
    for p, seed in ((512, 5), (96, 6)):
        X = _x(8192, p, seed)
        ref, _ = _run(F, X, "0", monkeypatch)
        monkeypatch.setenv("FH_BSR_CHAIN", "1")
        Y = torch.empty_like(X)
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        rc = L.fh_faust_multiply_noalloc(F.h, b"N", C.c_void_p(X.data_ptr()), p, C.c_void_p(Y.data_ptr()), st)
        assert rc == _lib.FH_OK, L.fh_last_error()
        torch.cuda.synchronize()
        assert torch.equal(ref, Y.view(torch.int16))
