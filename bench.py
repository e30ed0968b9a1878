"""Benchmark of the Faust x dense hot path (and PALM4MSA) on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu-baseline] [--no-palm]
               [--no-butterfly] [--no-notebook] [--hadamard]

Prints ONE JSON line.  `value` is the algorithmic HBM throughput (GB/s) of
one step = one Faust x dense product (all factors) with inputs resident in
HBM.  `roofline` reports the dominant kernel's algorithmic bytes per launch
divided by its average launch duration inside the graph-timed steps (HIP
events on the launch stream around the replay; the layout-conversion time
per step, from the library's per-factor event timers, is subtracted).  `cpu_baseline` is the
oracle (CPU FP64 restatement of the reference) timed on this host on a
bounded sample of the same workload.  `palm` times, on 512 targets, the
tuned batched PALM4MSA (FP32 + FP64 frozen-support refinement, HIP graphs,
autotuned tiles) until every target has reached, FP64-verified, the error
its FP64 PALM4MSA reaches in 220 iterations, against the plain FP64 GPU form
and the batched FP64 form.  `butterfly` times the butterfly factorization
(fact.butterfly, host matrix in, Faust out) at orders 1024 / 4096 / 16384 on
Gaussian matrices, next to the oracle at order 1024.  `notebook` times
F.toarray() and F.norm(2) on the 1024x1024 x10 Fausts whose GPU timings the
reference publishes.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="chain_bsr_bf16")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-palm", action="store_true",
                    help="skip the batched PALM4MSA comparison (512 targets, tuned vs plain FP64)")
    ap.add_argument("--palm-plain-all", action="store_true",
                    help="time the plain FP64 PALM4MSA on all 512 targets (about 6 minutes) instead of one per size")
    ap.add_argument("--no-butterfly", action="store_true", help="skip the butterfly factorization timings")
    ap.add_argument("--no-notebook", action="store_true",
                    help="skip F.toarray() / F.norm(2) on the reference notebook's 1024x1024 x10 Fausts")
    ap.add_argument("--butterfly-init", type=int, default=0, metavar="NHOLD",
                    help="also run PALM4MSA on NHOLD 256x256 butterfly targets from the reference start and from "
                         "the butterfly factorization")
    ap.add_argument("--hadamard", action="store_true",
                    help="also time hierarchical PALM4MSA on the 1024 Hadamard matrix, 100 iterations per level")
    ap.add_argument("--hadamard-its", type=int, default=100, metavar="ITS",
                    help="iterations per level of the --hadamard run (fewer for profiler runs)")
    ap.add_argument("--warmstart", type=int, default=0, metavar="NTRAIN",
                    help="also run the learned warm-start experiment (256x256 butterfly targets) with NTRAIN "
                         "training targets and 1000 held-out ones")
    return ap.parse_args()


# ----------------------------------------------------------------- workloads
def workload_chain_csr_f64():
    """5 CSR factors 1024x1024, exactly 8 distinct uniform columns per row,
    N(0,1) FP64; X 1024x64 N(0,1); one step = F@X and F.T@X."""
    import faust_amd
    n, J, nnz, p = 1024, 5, 8, 64
    facs = [faust_amd.rand_csr(n, n, nnz, np.random.default_rng(0xFA057 + f)) for f in range(J)]
    X = np.asfortranarray(np.random.default_rng(0xFA057 + 100).standard_normal((n, p)))
    per_factor = nnz * n * (8 + 4) + (n + 1) * 4 + n * p * 8 + n * p * 8
    chain = J * (nnz * n * (8 + 4) + (n + 1) * 4) + 2 * n * p * 8  # factors, X, Y of one product
    return dict(name="faust_csr_fp64_1024x1024_J5_nnz8_X1024x64", facs=facs, X=X, dtype="f64", ops=("N", "T"),
                bytes_per_launch=per_factor, launches_per_step=2 * J, bytes_per_step=2 * J * per_factor,
                chain_bytes_per_product=chain,
                flops_per_step=2 * J * 2 * n * nnz * p, kernel="k_csr_spmm_f64")


def _bsr_factors(seed, J=6, n=8192, bs=32, bpr=4):
    from faust_amd.bsr import rand_bsr
    return [rand_bsr(n, n, bs, bpr, np.random.default_rng(seed + f), scale=1 / np.sqrt(128)) for f in range(J)]


def _bsr_workload(values, act="bf16"):
    """6 BSR factors 8192x8192, 32x32 blocks, 4 distinct uniform block columns
    per block-row, values N(0, 1/128); X 8192x512 N(0,1); FP32 accumulation
    on MFMA.  values "bf16": BF16 blocks, BF16 activations; "f32" with act
    "bf16": FP32 blocks (kept as a BF16 hi/lo pair on the BF16 MFMA), BF16
    activations; "f32" with act "f32": FP32 blocks on the FP32 MFMA with FP32
    activations (the reference's MatBSR<float> precision)."""
    n, J, bs, bpr, p = 8192, 6, 32, 4, 512
    facs = _bsr_factors(0xB5B, J, n, bs, bpr)
    X = np.asfortranarray(np.random.default_rng(0xB5B + 100).standard_normal((n, p)))
    vbytes = 2 if values == "bf16" else 4
    abytes = 2 if act == "bf16" else 4
    nb = (n // bs) * bpr
    per_factor = nb * bs * bs * vbytes + nb * 4 + (n // bs + 1) * 4 + n * p * abytes * 2
    if act == "f32":
        kernel = "k_bsr_f32s"
    else:
        kernel = "k_bsr_shared<false,*,*>" if values == "bf16" else "k_bsr_shared<true,*,*>"
    return dict(name=f"faust_bsr_{values}_8192x8192_J6_b32x32_4blk_X8192x512_{act}", facs=facs, X=X,
                dtype=act, fac_dtype=("bf16" if values == "bf16" else np.float32),
                act_dtype=(np.float32 if act == "f32" else None), ops=("N",),
                bytes_per_launch=per_factor, launches_per_step=J, bytes_per_step=J * per_factor,
                flops_per_step=J * 2 * nb * bs * bs * p, kernel=kernel, cpu_reps_cap=40)


def workload_chain_bsr_bf16():
    return _bsr_workload("bf16")


def workload_chain_bsr_f32():
    return _bsr_workload("f32")


def workload_chain_bsr_f32_act32():
    return _bsr_workload("f32", "f32")


def workload_chain_csr_f32():
    """6 CSR factors 16384x16384, exactly 12 distinct uniform columns per row,
    N(0,1) FP32; X 16384x256 FP32 N(0,1).  (The number of factors is not
    fixed by the config text; 6 as for the BSR Faust.)"""
    import faust_amd
    n, J, nnz, p = 16384, 6, 12, 256
    facs = [faust_amd.rand_csr(n, n, nnz, np.random.default_rng(0xC5F + f)) for f in range(J)]
    facs = [f.astype(np.float32).astype(np.float64) for f in facs]
    X = np.asfortranarray(np.random.default_rng(0xC5F + 100).standard_normal((n, p)).astype(np.float32)
                          .astype(np.float64))
    per_factor = nnz * n * (4 + 4) + (n + 1) * 4 + n * p * 4 + n * p * 4
    return dict(name="faust_csr_fp32_16384x16384_J6_nnz12_X16384x256", facs=facs, X=X, dtype="f32",
                fac_dtype=np.float32, ops=("N",), bytes_per_launch=per_factor, launches_per_step=J,
                bytes_per_step=J * per_factor, flops_per_step=J * 2 * n * nnz * p, kernel="k_csr_rm_f32<*,12,true>",
                cpu_reps_cap=3)


WORKLOADS = {"chain_csr_f64": workload_chain_csr_f64, "chain_bsr_bf16": workload_chain_bsr_bf16,
             "chain_bsr_f32": workload_chain_bsr_f32, "chain_bsr_f32_act32": workload_chain_bsr_f32_act32,
             "chain_csr_f32": workload_chain_csr_f32}


# ----------------------------------------------------------------- PALM4MSA
def palm_targets(nt=512, seed=0xC4):
    """512 square targets, n uniform in {64,128,256,512}, each the product of
    3 random sparse n x n factors (4 nnz per row at uniform distinct columns,
    N(0,1)) plus i.i.d. N(0, 1e-6) noise (standard deviation 1e-3)."""
    import scipy.sparse as sp
    import faust_amd
    rng = np.random.default_rng(seed)
    sizes = rng.choice([64, 128, 256, 512], size=nt)
    As = []
    for n in sizes:
        P = sp.identity(int(n), format="csr")
        for _ in range(3):
            P = P @ faust_amd.rand_csr(int(n), int(n), 4, rng)
        As.append(np.asfortranarray(P.toarray() + 1e-3 * rng.standard_normal((n, n))))
    return sizes, As


def _relerr(lam, facs, A):
    P = np.eye(A.shape[0])
    for f in facs:
        P = P @ f
    return float(np.linalg.norm(lam * P - A) / np.linalg.norm(A))


def palm_batched_vs_plain(cpu=True, plain_all=False):
    """Time to the same FP64-verified relative error, 512 targets (3 factors,
    sp k = 4n).  The error each target has to reach is the one its FP64
    PALM4MSA reaches in 220 iterations, taken from the batched FP64 run (the
    plain FP64 form's own error is checked against it on every target the
    plain form runs).  Tuned form: FP32 iterations on the batched grouped-GEMM
    engine (one HIP graph per iteration, autotuned tiles; at most 400), then
    FP64 frozen-support sweeps, with an FP64 error check every 10 iterations;
    its time runs until every target has reached its error (set-up
    included).  The FP32 iterations end 0.15 % above the FP64 error at 220
    iterations (median), so the tuned form runs a few more of them.
    Plain form: the oracle-exact FP64 GPU path, one target at a time, 220
    iterations; timed on every target with plain_all, else on one target of
    each size and extrapolated by the size counts."""
    import torch
    from faust_amd import fact, factparams
    from faust_amd.batched import palm4msa_batched
    sizes, As = palm_targets()
    ks = [4 * int(n) for n in sizes]
    torch.cuda.synchronize()
    t = time.perf_counter()
    # the FP64 form with the oracle's cold power-iteration starts: its errors
    # are the plain form's (to 1e-9 on sampled targets); warm starts move the
    # 220-iteration error by up to 1 %
    b64 = palm4msa_batched(As, 3, k=ks, dtype="f64", num_its=220, refine_its=0, warm_power=False,
                           return_factors=False)
    t_b64 = time.perf_counter() - t
    goal = np.asarray(b64.relerr, dtype=np.float64) * (1 + 1e-9)
    torch.cuda.synchronize()
    t = time.perf_counter()
    # FP32 iterations (at most 400) until every target is there, then, for
    # targets still short of it, FP64 frozen-support sweeps
    tuned = palm4msa_batched(As, 3, k=ks, dtype="f32", num_its=400, refine_its=20, return_factors=False,
                             check_every=10, target_err=goal)
    t_tuned_wall = time.perf_counter() - t
    reached = tuned.iters_to_target >= 0
    t_tuned = (tuned.times_ms[0] + tuned.times_ms[4]) / 1e3 if tuned.times_ms[4] >= 0 else None
    idx = range(len(As)) if plain_all else [int(np.flatnonzero(sizes == n)[0]) for n in (64, 128, 256, 512)]
    plain_s, plain_err, by_size = 0.0, {}, {}
    for cnt, i in enumerate(idx):
        if plain_all and cnt % 32 == 0:
            print(f"plain FP64 PALM4MSA: target {cnt} of {len(idx)}, {plain_s:.1f} s so far", file=sys.stderr,
                  flush=True)
        n = int(sizes[i])
        p = factparams.ParamsPalm4MSA([factparams.ConstraintInt("sp", n, n, 4 * n)] * 3,
                                      factparams.StoppingCriterion(num_its=220))
        t = time.perf_counter()
        r = fact.palm4msa(As[i], p, ret_raw=True)
        dt = time.perf_counter() - t
        plain_err[i] = _relerr(r.lam, r.dense_factors(), As[i])
        plain_s += dt
        by_size.setdefault(n, []).append(dt)
    if not plain_all:
        plain_s = sum(by_size[n][0] * int(np.sum(sizes == n)) for n in by_size)
    dev = [abs(plain_err[i] - b64.relerr[i]) / plain_err[i] for i in plain_err]
    out = {"workload": "batched_palm4msa_512_targets_n64-512_3facs_sp4n_time_to_fp64_220it_error",
           "targets": len(As), "check_every": 10,
           "tuned_s_to_error": None if t_tuned is None else round(t_tuned, 3),
           "tuned_targets_reached": int(reached.sum()), "tuned_wall_s": round(t_tuned_wall, 3),
           "tuned_phase_ms": [round(x, 1) for x in tuned.times_ms[:5]],
           "tuned_iters_to_error_median": float(np.median(tuned.iters_to_target[reached])) if reached.any() else None,
           "batched_fp64_s": round(t_b64, 3),
           "plain_fp64_gpu_s": round(plain_s, 2), "plain_timed_targets": len(plain_err),
           "plain_measured": "all targets" if plain_all else "one target per size, extrapolated by the size counts",
           "plain_s_per_target_by_size": {str(n): round(float(np.mean(v)), 4) for n, v in sorted(by_size.items())},
           "plain_vs_batched_fp64_error_max_rel_diff": float(max(dev)),
           "median_relerr_goal": float(np.median(goal)),
           "median_relerr_tuned": float(np.median(tuned.relerr))}
    if t_tuned is not None:
        out["speedup_tuned_vs_plain"] = round(plain_s / t_tuned, 1)
        out["speedup_tuned_vs_batched_fp64"] = round(t_b64 / t_tuned, 2)
    if cpu:
        import oracle
        i = int(np.flatnonzero(sizes == 128)[0])
        t = time.perf_counter()
        oracle.palm4msa(As[i], [("sp", 4 * 128, 128, 128)] * 3, num_its=220)
        dt = time.perf_counter() - t
        est = sum(dt * (n / 128.0) ** 3 * int(np.sum(sizes == n)) for n in (64, 128, 256, 512))
        out["cpu_baseline"] = {"value": round(est, 1), "unit": "s (512 targets, n^3-extrapolated)",
                               "cores": len(os.sched_getaffinity(0)), "kind": "port",
                               "sample": f"oracle FP64 palm4msa on one n=128 target, 220 its: {dt:.3f} s"}
    return out


def palm_hadamard(its=100):
    """Hierarchical PALM4MSA of the 1024 Hadamard matrix into 10 factors
    (splincol, `its` iterations per level, plain FP64 GPU form)."""
    import oracle
    from faust_amd import fact, factparams
    H = oracle.hadamard(1024)
    t = time.perf_counter()
    res = fact.hierarchical(H, factparams.ParamsHierarchicalWHT(10, num_its=its), ret_raw=True)
    dt = time.perf_counter() - t
    return {"workload": f"hierarchical_hadamard_1024_10facs_splincol_{its}its", "seconds": round(dt, 3),
            "relerr": _relerr(res.lam, res.dense_factors(), H)}


def butterfly_measure(cpu=True):
    """fact.butterfly for 'bbtree' and 'right' on N(0,1) matrices: `gpu_ms`
    with the matrix already in HBM (a CUDA tensor, fh_butterfly_device),
    `gpu_ms_host_input` with the n x n FP64 host matrix uploaded by the call
    (fh_butterfly); both until the Faust is assembled, each the second of two
    calls.  Every run is checked on an exact butterfly of the same order
    (recovered to <= 1e-12)."""
    import torch
    from faust_amd import Faust, fact
    from faust_amd.butterfly import rand_butterfly_factors
    out = {"workload": "butterfly_factorization_gaussian_fp64", "gpu_ms": {}, "gpu_ms_host_input": {},
           "exact_relerr_max": 0.0}

    def timed(X, t):
        for _ in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            F = fact.butterfly(X, type=t)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            del F
        return round(dt * 1e3, 2)

    for n in (1024, 4096, 16384):
        rng = np.random.default_rng(n)
        M = np.asfortranarray(rng.standard_normal((n, n)))
        Md = torch.from_numpy(np.ascontiguousarray(M)).cuda()
        B = Faust(rand_butterfly_factors(n, rng=rng)).toarray()
        for t in ("bbtree", "right"):
            out["gpu_ms"][f"{t}_{n}"] = timed(Md, t)
            out["gpu_ms_host_input"][f"{t}_{n}"] = timed(M, t)
            G = fact.butterfly(B, type=t)
            e = float(np.linalg.norm(G.toarray() - B) / np.linalg.norm(B))
            out["exact_relerr_max"] = max(out["exact_relerr_max"], e)
            del G
        del Md
        print(f"butterfly n={n} done", file=sys.stderr, flush=True)
    if cpu:
        from oracle import butterfly as ob
        M = np.random.default_rng(1024).standard_normal((1024, 1024))
        t0 = time.perf_counter()
        ob.butterfly(M, ob.BALANCED)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(dt * 1e3, 1), "unit": "ms (bbtree, order 1024)", "cores": 1,
                               "kind": "port", "sample": "oracle bbtree factorization of one 1024 x 1024 N(0,1) "
                               "matrix (a sequential loop of numpy LAPACK SVDs, one per block)"}
        out["speedup_bbtree_1024"] = round(dt * 1e3 / out["gpu_ms"]["bbtree_1024"], 1)
    return out


def notebook_cases():
    """The reference's published GPU timings (gen_doc/using_gpu_pyfaust.ipynb,
    Tesla V100) of F.toarray() and F.norm(2) on pyfaust.rand(1024, 1024,
    num_factors=10, fac_type='dense') and fac_type='sparse', density=.2, FP64:
    the same calls on Fausts of the same shapes, median of 20 after 3 warm-ups,
    host array out as in pyfaust."""
    import torch
    import faust_amd
    out = {}
    for name, kw in (("dense", dict(fac_type="dense")), ("sparse_d0.2", dict(fac_type="sparse", density=0.2))):
        F = faust_amd.rand(1024, 1024, num_factors=10, seed=0, **kw)
        for what, fn in (("toarray_ms", F.toarray), ("norm2_ms", lambda: F.norm(2))):
            ts = []
            for _ in range(23):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                fn()
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            out[f"{name}_{what}"] = round(float(np.median(ts[3:])) * 1e3, 3)
    out["reference_v100_ms"] = {"dense_toarray_ms": 6.82, "dense_norm2_ms": 6.85, "sparse_d0.2_toarray_ms": 13.8,
                                "sparse_d0.2_norm2_ms": 9.86}
    return out


# ----------------------------------------------------------------- helpers
def accuracy(wl, F, Xd, out_N, ncols=16):
    """Error of the benchmarked product against the FP64 oracle on its first
    `ncols` columns (outside the timed region): `same_inputs` feeds the oracle
    the values the GPU holds (factor values and X after their rounding to the
    stored precision), so it measures the arithmetic (activation roundings,
    FP32 accumulation); `fp64_inputs` feeds it the unrounded FP64 factors and
    X, the whole error of the reduced-precision form.  FP64 workloads are also
    compared bit for bit."""
    import torch
    import oracle
    k = min(ncols, wl["X"].shape[1])
    Y = out_N[:k].to(torch.float64).cpu().numpy().T  # out_N holds the output column-major
    Xg = Xd[:k].to(torch.float64).cpu().numpy().T
    held = [f.to_host() for f in F._facs]

    def err(facs, X):
        ref = oracle.chain_multiply(facs, np.asfortranarray(X), "N")
        return {"rel_fro": float(np.linalg.norm(Y - ref) / np.linalg.norm(ref)),
                "max_abs_over_max_ref": float(np.abs(Y - ref).max() / np.abs(ref).max())}, ref

    same, ref_same = err(held, Xg)
    full, _ = err(wl["facs"], wl["X"][:, :k])
    out = {"columns": k, "same_inputs": same, "fp64_inputs": full}
    if wl["dtype"] == "f64":
        out["bitexact_vs_oracle"] = bool(np.array_equal(Y, ref_same))
    return out


def measured_traffic(wl):
    """HBM-side bytes per launch of the dominant kernel from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs recorded in profiles/traffic.json (each
    counter scaled by its reading on a plain copy of known size); None when
    no record matches this workload and kernel."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(wl["name"])
    except (OSError, ValueError):
        return None
    if not rec or rec.get("kernel") != wl["kernel"]:
        return None
    return int(rec["fetch_kib"] * 1024 * rec["fetch_scale"] + rec["write_kib"] * 1024 * rec["write_scale"])



def plain_copy_rate(nbytes, reps):
    """A plain device copy moving `nbytes` per copy (nbytes / 2 read, nbytes / 2
    written), `reps` dependent copies captured in one HIP graph and replayed,
    timed with HIP events: the rate a bandwidth-bound launch of that size gets
    on this GPU with no arithmetic and no gathers.  Returns us per copy, GB/s."""
    import torch
    n = max(1, nbytes // 8)
    a = torch.empty(n, dtype=torch.float32, device="cuda").normal_()
    b = torch.empty_like(a)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(reps):
            b.copy_(a)
    g.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    g.replay()
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    return round(us, 3), round(8 * n / (us * 1e-6) / 1e9, 1)


def cpu_baseline(wl, budget_s):
    import oracle
    nthreads = len(os.sched_getaffinity(0))
    X = wl["X"]
    # one warm-up rep, then as many reps as fit in the budget
    t0 = time.perf_counter()
    for op in wl["ops"]:
        oracle.chain_multiply(wl["facs"], X, op, nthreads=nthreads)
    one = time.perf_counter() - t0
    reps = max(1, min(wl.get("cpu_reps_cap", 200), int(budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(reps):
        for op in wl["ops"]:
            oracle.chain_multiply(wl["facs"], X, op, nthreads=nthreads)
    dt = (time.perf_counter() - t0) / reps
    return {"value": wl["bytes_per_step"] / dt / 1e9, "unit": "GB/s", "cores": nthreads, "kind": "port",
            "sample": f"{reps} steps of the same workload ({'+'.join(wl['ops'])}), oracle FP64 chain product, "
                      f"OpenMP over {nthreads} threads, {dt * 1e3:.3f} ms/step"}


def main():
    args = _parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 (multi-GPU is design-only)"}))
        sys.exit(2)
    import torch
    import faust_amd
    from faust_amd import _lib
    from faust_amd.faust import colmajor_to_device

    torch.cuda.set_device(0)
    wl = WORKLOADS[args.workload]()
    tdt = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}[wl["dtype"]]
    F = faust_amd.Faust(wl["facs"], dtype=wl.get("fac_dtype", np.float64), act_dtype=wl.get("act_dtype"))
    Xd = colmajor_to_device(wl["X"], tdt)
    FT = F.T
    outs = {op: torch.empty((Xd.shape[0], (F if op == "N" else FT).shape[0]), dtype=tdt, device="cuda")
            for op in wl["ops"]}

    def step():
        for op in wl["ops"]:
            (F if op == "N" else FT).matmul_device(Xd, out=outs[op])

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    L = _lib.lib()
    # FP64 CSR chains: the autotuner keeps, per product, the one-launch chain
    # or one launch per factor.  A one-launch product's algorithmic bytes are
    # its factors, X and Y (the intermediates stay in LDS); a per-factor
    # launch's are the workload's per-factor bytes.  `value` keeps the
    # factor-by-factor byte count of the workload.
    if wl["dtype"] == "f64":
        n_l = C.c_int()
        per_op, total, nbytes = {}, 0, 0
        for op in wl["ops"]:
            _lib.check(L.fh_faust_last_launches((F if op == "N" else FT).h, C.byref(n_l)))
            per_op[op] = n_l.value
            total += n_l.value
            nbytes += wl["chain_bytes_per_product"] if n_l.value == 1 else n_l.value * wl["bytes_per_launch"]
        if total != wl["launches_per_step"]:
            wl["launches_per_step"] = total
            wl["bytes_per_launch"] = nbytes // total
            wl["kernel"] = " + ".join(f"{op}: " + ("k_chain_pf_f64 (one launch)" if n == 1 else f"k_csr_spmm_f64 x{n}")
                                      for op, n in per_op.items())
    # timed region, two ways over the same K steps: (a) eager launches on the
    # stream, queued behind a GPU spin so that host enqueue time is off the
    # clock; (b) the K steps captured once in a HIP graph and replayed.  The
    # faster is reported (a graph replay adds a fixed cost per kernel node on
    # this runtime); both are in `config`.
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    torch.cuda._sleep(50_000_000)
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    wall_eager = time.perf_counter() - t0
    ms_eager = e0.elapsed_time(e1) / args.steps
    ms_graph = None
    wall_graph = None
    try:
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(args.steps):
                step()
        g.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        wall_graph = time.perf_counter() - t0
        ms_graph = e0.elapsed_time(e1) / args.steps
    except Exception:
        torch.cuda.synchronize()
    graphed = ms_graph is not None and ms_graph < ms_eager
    ms_step = ms_graph if graphed else ms_eager
    wall = wall_graph if graphed else wall_eager
    value = wl["bytes_per_step"] / (ms_step * 1e-3) / 1e9

    # per-factor HIP events on the launch stream around each factor's kernel,
    # over a second run of the same steps
    J = F.numfactors()
    for G in {id(F): F, id(FT): FT}.values():
        _lib.check(L.fh_faust_enable_timers(G.h, 1))
    torch.cuda._sleep(50_000_000)  # let the host queue every launch before the GPU starts
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = np.zeros(J + 1)
    cnt = np.zeros(J + 1, dtype=np.int64)
    for G in {id(F): F, id(FT): FT}.values():
        m1, c1 = np.zeros(J + 1), np.zeros(J + 1, dtype=np.int64)
        _lib.check(L.fh_faust_timers(G.h, m1.ctypes.data_as(C.c_void_p), c1.ctypes.data_as(C.c_void_p)))
        _lib.check(L.fh_faust_enable_timers(G.h, 0))
        ms += m1
        cnt += c1
    k_event_ms = ms[:J].sum() / max(cnt[:J].sum(), 1)
    conv_ms_step = ms[J] / max(args.steps, 1)
    # average launch duration of the factor kernel inside the timed steps:
    # (step time - layout conversions) / launches per step; the per-factor
    # event pairs (which add ~2 us of event overhead per launch) are reported
    # beside it
    n_ops = len(wl["ops"])
    k_ms = (ms_step - n_ops * conv_ms_step) / wl["launches_per_step"]
    achieved = wl["bytes_per_launch"] / (k_ms * 1e-3) / 1e9
    traffic = measured_traffic(wl)
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": wl["kernel"],
            "avg_launch_us": round(k_ms * 1e3, 3), "bytes_per_launch": wl["bytes_per_launch"],
            "per_factor_event_avg_launch_us": round(k_event_ms * 1e3, 3), "launches_timed": int(cnt[:J].sum()),
            "launches_per_step": wl["launches_per_step"],
            "layout_conversion_us_per_step": round(conv_ms_step * 1e3, 3)}
    # context for `frac`: a plain copy moving the same bytes per launch, and a
    # 256 MiB copy (the asymptotic copy rate)
    cu, cr = plain_copy_rate(wl["bytes_per_launch"], 200)
    lu, lr = plain_copy_rate(1 << 29, 10)
    roof["plain_copy"] = {"bytes_per_copy": 8 * max(1, wl["bytes_per_launch"] // 8), "us_per_copy": cu,
                          "GB/s": cr, "frac": round(cr / HBM_PEAK_GBS, 4),
                          "GB/s_256MiB_copy": lr, "frac_256MiB_copy": round(lr / HBM_PEAK_GBS, 4)}

    out = {"metric": "Faust·X HBM GB/s (% peak); PALM4MSA time-to-rel-err, tuned vs plain FP64",
           "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms_step, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": wl["dtype"], "data": "synthetic (seeded N(0,1) values, uniform random supports)",
           "config": {"workload": wl["name"], "ops": list(wl["ops"]), "bytes_per_step": wl["bytes_per_step"],
                      "frac_of_hbm_peak": round(value / HBM_PEAK_GBS, 4), "wall_s": round(wall, 4),
                      "timed_as": "hip graph replay" if graphed else "eager stream launches",
                      "ms_per_step_eager": round(ms_eager, 5),
                      "ms_per_step_graph": None if ms_graph is None else round(ms_graph, 5)},
           "roofline": roof}
    if "N" in wl["ops"]:
        out["accuracy_vs_fp64_oracle"] = accuracy(wl, F, Xd, outs["N"])
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_seconds)
    if not args.no_palm:
        out["palm"] = palm_batched_vs_plain(cpu=not args.no_cpu_baseline, plain_all=args.palm_plain_all)
    if not args.no_butterfly:
        out["butterfly"] = butterfly_measure(cpu=not args.no_cpu_baseline)
    if args.hadamard:
        out["palm_hadamard"] = palm_hadamard(args.hadamard_its)
    if not args.no_notebook:
        out["notebook"] = notebook_cases()
    if args.butterfly_init:
        from faust_amd import warmstart
        out["butterfly_init"] = warmstart.butterfly_init_experiment(
            nhold=args.butterfly_init, log=lambda m: print(m, file=sys.stderr, flush=True))
    if args.warmstart:
        from faust_amd import warmstart
        out["warmstart"] = warmstart.experiment(ntrain=args.warmstart, log=lambda m: print(m, file=sys.stderr))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
