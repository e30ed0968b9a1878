"""Benchmark of the Faust x dense hot path (and PALM4MSA) on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu-baseline] [--palm]

Prints ONE JSON line.  `value` is the algorithmic HBM throughput (GB/s) of
one step = one Faust x dense product (all factors) with inputs resident in
HBM.  `roofline` reports the dominant kernel's algorithmic bytes per launch
divided by its average launch duration (HIP events around a HIP-graph replay
of back-to-back launches on the launch stream).  `cpu_baseline` is the
oracle (CPU FP64 restatement of the reference) timed on this host on a
bounded sample of the same workload.
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
    ap.add_argument("--workload", default="chain_csr_f64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--palm", action="store_true", help="also time hierarchical PALM4MSA (Hadamard 1024)")
    ap.add_argument("--kernel-reps", type=int, default=200)
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
    return dict(name="faust_csr_fp64_1024x1024_J5_nnz8_X1024x64", facs=facs, X=X, dtype="f64", ops=("N", "T"),
                bytes_per_launch=per_factor, launches_per_step=2 * J, bytes_per_step=2 * J * per_factor,
                flops_per_step=2 * J * 2 * n * nnz * p, kernel="k_csr_spmm_f64")


WORKLOADS = {"chain_csr_f64": workload_chain_csr_f64}


# ----------------------------------------------------------------- helpers
def _graph_time(fn, reps, warm=3):
    """Average time (ms) of fn() replayed from a captured HIP graph, timed with
    events on the capture stream.  Falls back to an eager loop."""
    import torch
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for _ in range(warm):
            fn()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    try:
        with torch.cuda.graph(g, stream=s):
            for _ in range(reps):
                fn()
        graphed = True
    except Exception:
        graphed = False
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(s):
        if graphed:
            g.replay()  # warm replay
            torch.cuda.current_stream().synchronize()
            e0.record()
            g.replay()
            e1.record()
        else:
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps, graphed


def cpu_baseline(wl, budget_s):
    import oracle
    nthreads = len(os.sched_getaffinity(0))
    X = wl["X"]
    # one warm-up rep, then as many reps as fit in the budget
    t0 = time.perf_counter()
    for op in wl["ops"]:
        oracle.chain_multiply(wl["facs"], X, op, nthreads=nthreads)
    one = time.perf_counter() - t0
    reps = max(1, min(200, int(budget_s / max(one, 1e-6))))
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
    F = faust_amd.Faust(wl["facs"], dtype={"f64": np.float64, "f32": np.float32, "bf16": "bf16"}[wl["dtype"]])
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
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_time(e1) / args.steps
    value = wl["bytes_per_step"] / (ms_step * 1e-3) / 1e9

    # dominant kernel: one factor product, back-to-back launches in a HIP graph
    f0 = F._facs[0]
    p = Xd.shape[0]
    r, c = f0.shape
    xin = torch.randn((p, c), dtype=tdt, device="cuda")
    yout = torch.empty((p, r), dtype=tdt, device="cuda")
    L = _lib.lib()
    one, zero = (C.c_double(1.0), C.c_double(0.0)) if wl["dtype"] == "f64" else (C.c_float(1.0), C.c_float(0.0))

    def launch():
        _lib.check(L.fh_mat_mul_dense(f0.h, b"N", C.c_void_p(xin.data_ptr()), c, p, C.c_void_p(yout.data_ptr()), r,
                                      C.byref(one), C.byref(zero),
                                      C.c_void_p(torch.cuda.current_stream().cuda_stream)))

    k_ms, graphed = _graph_time(launch, args.kernel_reps)
    achieved = wl["bytes_per_launch"] / (k_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": wl["kernel"],
            "avg_launch_us": round(k_ms * 1e3, 3), "graph_timed": graphed}

    out = {"metric": "Faust·X HBM GB/s (% peak); PALM4MSA time-to-rel-err, tuned vs plain FP64",
           "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms_step, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": wl["dtype"], "data": "synthetic (seeded N(0,1) values, uniform random supports)",
           "config": {"workload": wl["name"], "ops": list(wl["ops"]), "bytes_per_step": wl["bytes_per_step"],
                      "frac_of_hbm_peak": round(value / HBM_PEAK_GBS, 4), "wall_s": round(wall, 4)},
           "roofline": roof}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
