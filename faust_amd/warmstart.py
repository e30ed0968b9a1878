"""Learned warm start for butterfly PALM4MSA.

Targets are products of L = log2(n) butterfly factors; factor j couples
indices i and i ^ 2^j through a 2x2 block s * [[cos t, -sin t], [sin t, cos t]]
(t ~ U(0, 2 pi), s ~ U(0.5, 2)), plus i.i.d. Gaussian noise.  PALM4MSA
factorizes them with `supp` projectors on the butterfly supports.

An MLP (n^2 -> 512 -> 512 -> 2 n L, GELU; BF16 weights on MFMA, FP32
accumulation; libfaust_hip's own forward/backward/Adam kernels) is trained
on factorizations produced by the plain FP64 PALM4MSA and predicts the 2 n L
nonzero values of the factors.  PALM4MSA starts from the prediction and a
run falls back to the reference initialisation when its final error is
worse than the plain run's.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


def butterfly_supports(n):
    """0/1 masks of the L butterfly factors (row i: columns i and i ^ 2^j)."""
    L = int(np.log2(n))
    masks = []
    i = np.arange(n)
    for j in range(L):
        m = np.zeros((n, n))
        m[i, i] = 1.0
        m[i, i ^ (1 << j)] = 1.0
        masks.append(np.asfortranarray(m))
    return masks


def random_butterfly_factor(n, j, rng):
    B = np.zeros((n, n))
    for i in range(n):
        if i & (1 << j):
            continue
        i2 = i | (1 << j)
        t = rng.uniform(0.0, 2 * np.pi)
        s = rng.uniform(0.5, 2.0)
        c, sn = s * np.cos(t), s * np.sin(t)
        B[i, i], B[i, i2], B[i2, i], B[i2, i2] = c, -sn, sn, c
    return B


def butterfly_target(n, rng, noise=1e-6):
    """A = B_0 B_1 ... B_{L-1} + N(0, noise^2) entries."""
    return butterfly_sample(n, rng, noise)[0]


def butterfly_sample(n, rng, noise=1e-6):
    """(A, [B_0 .. B_{L-1}]): a target and the factors it was made from."""
    L = int(np.log2(n))
    i = np.arange(n)
    facs = []
    for j in range(L):
        lo = i[(i & (1 << j)) == 0]
        hi = lo | (1 << j)
        t = rng.uniform(0.0, 2 * np.pi, lo.size)
        s = rng.uniform(0.5, 2.0, lo.size)
        c, sn = s * np.cos(t), s * np.sin(t)
        B = np.zeros((n, n))
        B[lo, lo], B[lo, hi], B[hi, lo], B[hi, hi] = c, -sn, sn, c
        facs.append(B)
    P = np.eye(n)
    for B in facs:
        P = P @ B
    return np.asfortranarray(P + noise * rng.standard_normal((n, n))), facs


def pack_values(factors, masks):
    """The 2 n per-factor nonzero values, factor-major, row-major inside a factor."""
    out = []
    for F, M in zip(factors, masks):
        r, c = np.nonzero(M)  # row-major order (numpy nonzero scans rows)
        out.append(F[r, c])
    return np.concatenate(out)


def unpack_values(v, masks):
    out = []
    off = 0
    for M in masks:
        r, c = np.nonzero(M)
        F = np.zeros(M.shape)
        F[r, c] = v[off:off + len(r)]
        off += len(r)
        out.append(np.asfortranarray(F))
    return out


def identity_init(masks):
    return [np.asfortranarray(np.eye(M.shape[0]) * (np.diag(M) != 0)) for M in masks]


def canonicalize(factors):
    """Canonical representative of a butterfly factorization.

    F_0 ... F_{L-1} and F_0 D_0, D_0^{-1} F_1 D_1, ... (D_j diagonal,
    invertible) have the same product and the same supports, so raw factor
    values are not a function of the target.  Going left to right, every
    column i of F_j (j < L-1) is scaled to unit norm with a positive diagonal
    entry and row i of F_{j+1} takes the inverse scale; the last factor keeps
    what is left.  Then every factor is divided by its Frobenius norm, which
    goes into lambda.  Returns (lambda, factors) with the same product."""
    F = [np.array(f, dtype=np.float64) for f in factors]
    for j in range(len(F) - 1):
        nrm = np.linalg.norm(F[j], axis=0)
        sgn = np.where(np.diag(F[j]) < 0, -1.0, 1.0)
        d = np.where(nrm > 0, sgn / np.where(nrm > 0, nrm, 1.0), 1.0)
        F[j] = F[j] * d[None, :]
        F[j + 1] = F[j + 1] / d[:, None]
    lam = 1.0
    for j in range(len(F)):
        s = np.linalg.norm(F[j])
        lam *= s
        F[j] = np.asfortranarray(F[j] / s)
    return lam, F


class MLP:
    """The warm-start network on the GPU (fh_mlp_*)."""

    def __init__(self, n_in, n_h1=512, n_h2=512, n_out=4096, seed=0):
        self.dims = (n_in, n_h1, n_h2, n_out)
        self.h = C.c_void_p()
        check(lib().fh_mlp_create(n_in, n_h1, n_h2, n_out, seed, C.byref(self.h)))

    def __del__(self):
        lib_ = getattr(_lib, "_lib", None) if _lib is not None else None
        if getattr(self, "h", None) and self.h.value and lib_ is not None:
            lib_.fh_mlp_free(self.h)

    def _stream(self):
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def forward(self, Xd):
        """Xd: (B, n_in) float32 CUDA tensor -> (B, n_out) float32 CUDA tensor."""
        import torch
        Y = torch.empty((Xd.shape[0], self.dims[3]), dtype=torch.float32, device="cuda")
        check(lib().fh_mlp_forward(self.h, C.c_void_p(Xd.data_ptr()), Xd.shape[0], C.c_void_p(Y.data_ptr()),
                                   self._stream()))
        return Y

    def train_step(self, Xd, Td, lr):
        loss = C.c_float()
        check(lib().fh_mlp_train_step(self.h, C.c_void_p(Xd.data_ptr()), C.c_void_p(Td.data_ptr()), Xd.shape[0],
                                      float(lr), C.byref(loss), self._stream()))
        return loss.value

    def gradients(self, Xd, Td):
        loss = C.c_float()
        check(lib().fh_mlp_gradients(self.h, C.c_void_p(Xd.data_ptr()), C.c_void_p(Td.data_ptr()), Xd.shape[0],
                                     C.byref(loss), self._stream()))
        return loss.value, [self.get(l, what=1) for l in range(3)]

    def get(self, layer, what=0):
        K, N = self.dims[layer], self.dims[layer + 1]
        W = np.empty((K, N), dtype=np.float32)
        b = np.empty(N, dtype=np.float32)
        check(lib().fh_mlp_get(self.h, layer, what, W.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p)))
        return W, b

    def set(self, layer, W, b):
        W = np.ascontiguousarray(W, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        check(lib().fh_mlp_set(self.h, layer, W.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p)))


# ---------------------------------------------------------------- pipeline
def mlp_input(A):
    """Network input: the target scaled to unit RMS entry, flattened row-major."""
    A = np.asarray(A, dtype=np.float64)
    return (A * (A.shape[0] / np.linalg.norm(A))).astype(np.float32).ravel()


def label_scale(n):
    """Canonical factors have unit Frobenius norm over 2n entries: scale to O(1)."""
    return float(np.sqrt(2 * n))


def palm_labels(As, gens, masks, num_its=20, chunk=512):
    """Training labels made by the FP64 PALM4MSA itself: each target is
    factorized with supp projectors on the butterfly supports, starting from
    the canonical form of the factors it was generated from, and the result is
    brought to canonical form.  Returns (labels [N, 2nL] float32, lambdas,
    final FP64-verified relative errors)."""
    from .batched import palm4msa_batched
    n, L = As[0].shape[0], len(masks)
    labels, lams, errs = [], [], []
    for c0 in range(0, len(As), chunk):
        part = range(c0, min(len(As), c0 + chunk))
        can = [canonicalize(gens[b]) for b in part]
        r = palm4msa_batched([As[b] for b in part], L, supports=[masks] * len(part), init=[c[1] for c in can],
                             lambda_init=[c[0] for c in can], dtype="f64", num_its=num_its, refine_its=0)
        for q in range(len(part)):
            lam, F = canonicalize(r.factors[q])
            labels.append((pack_values(F, masks) * label_scale(n)).astype(np.float32))
            lams.append(lam * r.lambdas[q])
        errs.extend(r.relerr.tolist())
    return np.stack(labels), np.array(lams), np.array(errs)


def train_mlp(mlp, X, Y, epochs=10, batch=256, lr=1e-3, seed=0, log=None):
    """Adam on the MSE between network output and standardised labels (each
    output's training mean and standard deviation are kept on the network
    object and undone by `predicted_init`).  X, Y: float32 CUDA tensors
    [N, n_in], [N, n_out]."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    N = X.shape[0]
    mu = Y.mean(dim=0, keepdim=True)
    sd = Y.std(dim=0, keepdim=True).clamp_min(1e-6)
    mlp.label_mu, mlp.label_sd = mu, sd
    Y = ((Y - mu) / sd).contiguous()
    losses = []
    for ep in range(epochs):
        perm = torch.randperm(N, generator=g).to(X.device)
        tot, cnt = 0.0, 0
        lr_ep = lr * (0.5 ** (ep // max(1, epochs // 3)))
        for b0 in range(0, N - batch + 1, batch):
            idx = perm[b0:b0 + batch]
            xb, yb = X.index_select(0, idx).contiguous(), Y.index_select(0, idx).contiguous()
            tot += mlp.train_step(xb, yb, lr_ep)
            cnt += 1
        losses.append(tot / max(cnt, 1))
        if log:
            log(f"epoch {ep}: mse {losses[-1]:.5f} lr {lr_ep:.2e}")
    return losses


def predicted_init(mlp, As, masks, batch=256):
    """Initial factors and lambda for PALM4MSA from the network: unpack the
    prediction, normalise every factor like the supp projector does, and take
    the lambda that best fits the target (<A, P> / ||P||^2)."""
    import torch
    n = As[0].shape[0]
    out_f, out_l = [], []
    for b0 in range(0, len(As), batch):
        part = As[b0:b0 + batch]
        X = torch.from_numpy(np.stack([mlp_input(A) for A in part])).cuda()
        Y = mlp.forward(X)
        if getattr(mlp, "label_mu", None) is not None:
            Y = Y * mlp.label_sd + mlp.label_mu
        Y = Y.cpu().numpy().astype(np.float64) / label_scale(n)
        for A, y in zip(part, Y):
            F = [f / max(np.linalg.norm(f), 1e-300) for f in unpack_values(y, masks)]
            P = np.eye(n)
            for f in F:
                P = P @ f
            pp = float(np.sum(P * P))
            out_f.append([np.asfortranarray(f) for f in F])
            out_l.append(float(np.sum(A * P)) / pp if pp > 0 else 1.0)
    return out_f, out_l


def butterfly_init(As, masks):
    """Initial factors and lambda for PALM4MSA from the butterfly
    factorization (fact.butterfly, 'bbtree').  Factor j here couples the
    indices that differ in bit j, which is butterfly support L-1-j of
    faust_butterfly.hpp, so P A P (P the bit reversal, its own inverse) is a
    product in the reference's support order: its factors F_j give
    P F_j P on mask j, and the product is unchanged.  Each factor is
    normalised like the supp projector does and lambda fits the target
    (<A, P> / ||P||^2)."""
    from .butterfly import bitrev
    from .fact import butterfly
    n = As[0].shape[0]
    br = bitrev(n)
    out_f, out_l = [], []
    for A in As:
        F = butterfly(np.asfortranarray(A[np.ix_(br, br)]), type="bbtree")
        fs = []
        for j, M in enumerate(masks):
            f = F._facs[j].to_host().toarray()[np.ix_(br, br)]
            f = f * (M != 0)
            fs.append(np.asfortranarray(f / max(np.linalg.norm(f), 1e-300)))
        P = np.eye(n)
        for f in fs:
            P = P @ f
        pp = float(np.sum(P * P))
        out_f.append(fs)
        out_l.append(float(np.sum(A * P)) / pp if pp > 0 else 1.0)
    return out_f, out_l


class _Runs:
    """Concatenation of batched runs over consecutive chunks of targets."""

    def __init__(self, parts):
        self.relerr = np.concatenate([r.relerr for r in parts])
        self.iters_to_target = np.concatenate([r.iters_to_target for r in parts])
        # time until every target of a chunk reached the target error: summed
        # over chunks when all of them got there, else -1
        t = [r.times_ms[4] for r in parts]
        self.times_ms = np.array([0, 0, 0, 0, sum(t) if min(t) >= 0 else -1.0])


def compare_warm_start(As, masks, inits, lambdas, max_its=1000, check_every=25, target_err=1e-4, dtype="f64",
                       chunk=250, log=None):
    """PALM4MSA from the predicted initialisation and from the reference one
    (first factor 0, the others identity), same iteration budget, FP64-verified
    errors checked every `check_every` iterations.  A warm-started run whose
    final error is worse than the reference run's falls back to the latter.
    Targets go through the batched engine `chunk` at a time."""
    import time
    from .batched import palm4msa_batched
    L = len(masks)
    refs, warms = [], []
    t_ref = t_warm = 0.0
    for c0 in range(0, len(As), chunk):
        sl = slice(c0, min(len(As), c0 + chunk))
        part = As[sl]
        sup = [masks] * len(part)
        t = time.perf_counter()
        refs.append(palm4msa_batched(part, L, supports=sup, dtype=dtype, num_its=max_its, refine_its=0,
                                     check_every=check_every, target_err=target_err, return_factors=False))
        t_ref += time.perf_counter() - t
        t = time.perf_counter()
        warms.append(palm4msa_batched(part, L, supports=sup, init=inits[sl], lambda_init=lambdas[sl], dtype=dtype,
                                      num_its=max_its, refine_its=0, check_every=check_every,
                                      target_err=target_err, return_factors=False))
        t_warm += time.perf_counter() - t
        if log:
            log(f"PALM on held-out targets {sl.stop}/{len(As)}: reference {t_ref:.1f}s, warm {t_warm:.1f}s")
    ref, warm = _Runs(refs), _Runs(warms)
    # fall back when the warm run ended non-finite, or worse than a finite
    # reference run (a target whose FP32 iteration broke down ends with a
    # non-finite error instead of stopping the whole batch)
    ref_ok, warm_ok = np.isfinite(ref.relerr), np.isfinite(warm.relerr)
    fallback = ~warm_ok | (ref_ok & (warm.relerr > ref.relerr))
    final = np.where(fallback, ref.relerr, warm.relerr)
    it_w = np.where(fallback, ref.iters_to_target, warm.iters_to_target)

    def med(e):  # median with non-finite errors counted as the worst
        return float(np.median(np.where(np.isfinite(e), e, np.inf)))
    return {"targets": len(As), "max_its": max_its, "target_err": target_err, "dtype": dtype,
            "ref_reached": int((ref.iters_to_target > 0).sum()), "warm_reached": int((it_w > 0).sum()),
            "ref_median_its": float(np.median(ref.iters_to_target[ref.iters_to_target > 0]))
            if (ref.iters_to_target > 0).any() else None,
            "warm_median_its": float(np.median(it_w[it_w > 0])) if (it_w > 0).any() else None,
            "ref_median_err": med(ref.relerr), "warm_median_err": med(final),
            "ref_nonfinite": int((~ref_ok).sum()), "warm_nonfinite": int((~warm_ok).sum()),
            "warm_median_err_before_fallback": med(warm.relerr),
            "warm_reached_before_fallback": int((warm.iters_to_target > 0).sum()),
            "fallbacks": int(fallback.sum()), "ref_wall_s": round(t_ref, 3), "warm_wall_s": round(t_warm, 3),
            "ref_ms_to_all": float(ref.times_ms[4]), "warm_ms_to_all": float(warm.times_ms[4])}


def experiment(n=256, ntrain=16384, nhold=1000, epochs=40, label_its=5, max_its=1000, check_every=25,
               target_err=1e-4, noise=1e-6, hidden=512, dtype="f32", seed=0x5EED, log=print):
    """The learned-warm-start experiment end to end: butterfly targets, labels
    from FP64 PALM4MSA, MLP training on the GPU, then PALM4MSA on held-out
    targets from the predicted and from the reference initialisation, and
    from the butterfly factorization's factors (butterfly_init)."""
    import time
    import torch
    rng = np.random.default_rng(seed)
    masks = butterfly_supports(n)
    L = len(masks)
    t_gen = t_lab = 0.0
    Xs, Ys, lerr = [], [], []
    for c0 in range(0, ntrain, 512):  # training targets are streamed: only inputs and labels are kept
        t0 = time.perf_counter()
        part = [butterfly_sample(n, rng, noise) for _ in range(min(512, ntrain - c0))]
        t_gen += time.perf_counter() - t0
        t0 = time.perf_counter()
        Yc, _, ec = palm_labels([a for a, _ in part], [g for _, g in part], masks, num_its=label_its)
        t_lab += time.perf_counter() - t0
        Xs.append(np.stack([mlp_input(a) for a, _ in part]))
        Ys.append(Yc)
        lerr.extend(ec.tolist())
        log(f"training targets {c0 + len(part)}/{ntrain}: gen {t_gen:.1f}s labels {t_lab:.1f}s")
    hold = [butterfly_sample(n, rng, noise)[0] for _ in range(nhold)]
    log(f"targets: gen {t_gen:.1f}s labels {t_lab:.1f}s (median label relerr {np.median(lerr):.2e})")
    X = torch.from_numpy(np.concatenate(Xs)).cuda()
    Yd = torch.from_numpy(np.concatenate(Ys)).cuda()
    del Xs, Ys
    mlp = MLP(n * n, hidden, hidden, 2 * n * L, seed=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = train_mlp(mlp, X, Yd, epochs=epochs, batch=256, lr=1e-3, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del X, Yd
    t0 = time.perf_counter()
    inits, lam0 = predicted_init(mlp, hold, masks)
    t_pred = time.perf_counter() - t0
    log(f"predicted inits for {nhold} held-out targets in {t_pred:.1f}s")
    res = compare_warm_start(hold, masks, inits, lam0, max_its=max_its, check_every=check_every,
                             target_err=target_err, dtype=dtype, log=log)
    # the butterfly factorization as the initialiser
    t0 = time.perf_counter()
    binits, blam = butterfly_init(hold, masks)
    t_bf = time.perf_counter() - t0
    log(f"butterfly inits for {nhold} held-out targets in {t_bf:.1f}s")
    bres = compare_warm_start(hold, masks, binits, blam, max_its=max_its, check_every=check_every,
                              target_err=target_err, dtype=dtype, log=log)
    res["butterfly_init"] = {k: bres[k] for k in ("warm_reached", "warm_median_its", "warm_median_err",
                                                  "warm_reached_before_fallback", "fallbacks", "warm_nonfinite",
                                                  "warm_wall_s", "warm_ms_to_all")}
    res["butterfly_init"]["init_s"] = round(t_bf, 2)
    res.update({"n": n, "factors": L, "noise_std": noise, "train_targets": ntrain, "label_palm_its": label_its,
                "epochs": epochs, "final_train_mse": float(losses[-1]), "train_s": round(t_train, 2),
                "predict_s": round(t_pred, 2), "label_s": round(t_lab, 1), "mlp": [n * n, hidden, hidden, 2 * n * L]})
    return res


def butterfly_init_experiment(n=256, nhold=200, max_its=1000, check_every=25, target_err=1e-4, noise=1e-6,
                              dtype="f32", seed=0x5EED, log=print):
    """PALM4MSA on held-out butterfly targets from the reference start and
    from the butterfly factorization (butterfly_init), same iteration budget
    and checks as the learned-warm-start comparison."""
    import time
    rng = np.random.default_rng(seed)
    masks = butterfly_supports(n)
    hold = [butterfly_sample(n, rng, noise)[0] for _ in range(nhold)]
    t0 = time.perf_counter()
    inits, lam = butterfly_init(hold, masks)
    t_init = time.perf_counter() - t0
    log(f"butterfly inits for {nhold} targets in {t_init:.1f}s")
    res = compare_warm_start(hold, masks, inits, lam, max_its=max_its, check_every=check_every,
                             target_err=target_err, dtype=dtype, log=log)
    res.update({"n": n, "factors": len(masks), "noise_std": noise, "init": "butterfly factorization (bbtree)",
                "init_s": round(t_init, 2)})
    return res
