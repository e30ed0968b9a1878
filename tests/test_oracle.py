"""Pins the CPU oracle (oracle/) without the reference, which cannot be built
here (Eigen 3.4 absent) or imported (pyfaust absent):
closed forms, an independent library (numpy/scipy), hand-derived traces of
the libstdc++ heap selection, glibc rand, and exact Hadamard recovery."""
import ctypes
import ctypes.util
import math

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import faust_amd


def _rand_chain(rng, shapes, dens=0.3, integer=False):
    out = []
    for i, (r, c) in enumerate(shapes):
        a = rng.standard_normal((r, c)) if not integer else rng.integers(-3, 4, (r, c)).astype(float)
        a = a * (rng.random((r, c)) < dens)
        out.append(sp.csr_matrix(a) if i % 2 == 0 else np.asfortranarray(a))
    return out


def _dense(f):
    return f.toarray() if sp.issparse(f) else np.asarray(f)


def test_chain_product_vs_numpy():
    rng = np.random.default_rng(0)
    facs = _rand_chain(rng, [(40, 30), (30, 50), (50, 20), (20, 33)])
    P = _dense(facs[0])
    for f in facs[1:]:
        P = P @ _dense(f)
    X = rng.standard_normal((33, 7))
    Y = oracle.chain_multiply(facs, X, "N")
    assert np.linalg.norm(Y - P @ X) <= 1e-13 * np.linalg.norm(P @ X)
    X2 = rng.standard_normal((40, 5))
    Y2 = oracle.chain_multiply(facs, X2, "T")
    assert np.linalg.norm(Y2 - P.T @ X2) <= 1e-13 * np.linalg.norm(P.T @ X2)


def test_integer_chain_exact_identity_and_transpose():
    # unit/faust_mult.cpp.in: F * I == get_product() exactly on small integer factors
    rng = np.random.default_rng(1)
    facs = _rand_chain(rng, [(16, 16)] * 4, dens=0.4, integer=True)
    P = oracle.get_product(facs)
    assert np.array_equal(oracle.chain_multiply(facs, np.eye(16), "N"), P)
    P_np = _dense(facs[0])
    for f in facs[1:]:
        P_np = P_np @ _dense(f)
    assert np.array_equal(P, P_np)  # integers: every order is exact
    X = rng.integers(-2, 3, (16, 4)).astype(float)
    assert np.array_equal(oracle.chain_multiply(facs, X, "T"), (X.T @ P).T)


def test_bsr_chain_matches_dense():
    from faust_amd.bsr import rand_bsr
    rng = np.random.default_rng(2)
    B = rand_bsr(128, 96, 32, 2, rng)
    X = rng.standard_normal((96, 3))
    assert np.linalg.norm(oracle.chain_multiply([B], X, "N") - B.toarray() @ X) < 1e-12


def test_wht_factors_and_hadamard():
    n = 64
    H = oracle.hadamard(n, normalized=False)
    assert np.array_equal(H @ H.T, n * np.eye(n))
    P = oracle.get_product(faust_amd.wht_factors(n))
    assert np.abs(P - H / np.sqrt(n)).max() < 1e-14


def test_gemm_alpha_beta():
    rng = np.random.default_rng(3)
    A, B, C0 = rng.standard_normal((9, 7)), rng.standard_normal((9, 5)), rng.standard_normal((7, 5))
    got = oracle.gemm(A, B, "T", "N", 0.5, -1.0, C0)
    assert np.allclose(got, 0.5 * A.T @ B - C0, rtol=1e-14, atol=1e-14)
    Ai = rng.integers(-4, 5, (6, 6)).astype(float)
    assert np.array_equal(oracle.gemm(Ai, Ai, "N", "T"), Ai @ Ai.T)


def test_spectral_norm_closed_forms():
    d = np.array([3.0, -7.0, 2.0, 5.0])
    nrm, _ = oracle.spectral_norm([np.diag(d)])
    assert abs(nrm - 7.0) < 1e-6
    nrm, _ = oracle.spectral_norm(faust_amd.wht_factors(128))
    assert abs(nrm - 1.0) < 1e-9
    rng = np.random.default_rng(4)
    facs = _rand_chain(rng, [(30, 30), (30, 30)], dens=0.5)
    P = _dense(facs[0]) @ _dense(facs[1])
    nrm, _ = oracle.spectral_norm(facs, maxit=1000, thr=1e-14)
    assert abs(nrm - np.linalg.norm(P, 2)) < 1e-6 * np.linalg.norm(P, 2)
    # rectangular (F^H F chain) and a zero factor (is_zero -> 0)
    facs = [rng.standard_normal((20, 8)), rng.standard_normal((8, 5))]
    nrm, _ = oracle.spectral_norm(facs, maxit=1000, thr=1e-14)
    assert abs(nrm - np.linalg.norm(facs[0] @ facs[1], 2)) < 1e-6 * nrm
    assert oracle.spectral_norm([sp.csr_matrix((5, 5)), np.eye(5)])[0] == 0.0


def test_power_iteration_start_vector_is_glibc_eigen_random():
    # srand(0xF4+'u'+57); Eigen Random on [-1,1]: -1 + 2*rand()/RAND_MAX
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.srand(0xF4 + ord("u") + 57)
    RAND_MAX = 2147483647
    exp = np.array([-1.0 + (2.0 * libc.rand()) / RAND_MAX for _ in range(100)])
    assert np.array_equal(oracle.power_x0(100), exp)


def test_reductions():
    rng = np.random.default_rng(5)
    for n in (1, 100, 4096, 4097, 100000):
        x = rng.standard_normal(n)
        assert abs(oracle.reduce_sum(x) - math.fsum(x)) <= 1e-12 * max(1.0, np.abs(x).sum())
        assert abs(oracle.reduce_dot(x, x) - math.fsum(x * x)) <= 1e-12 * (x * x).sum()


# ---- projectors ---------------------------------------------------------
def test_prox_no_ties_vs_brute_force():
    rng = np.random.default_rng(6)
    M = rng.standard_normal((13, 11))
    k = 4
    got = oracle.prox("splin", M, k, normalized=False)
    for i in range(13):
        keep = np.argsort(-np.abs(M[i]))[:k]
        exp = np.zeros(11)
        exp[keep] = M[i, keep]
        assert np.array_equal(got[i], exp)
    got = oracle.prox("spcol", M, k, normalized=False)
    for j in range(11):
        keep = np.argsort(-np.abs(M[:, j]))[:k]
        assert set(np.nonzero(got[:, j])[0]) == set(keep)
    got = oracle.prox("sp", M, 20, normalized=True)
    keep = np.argsort(-np.abs(M.ravel(order="F")))[:20]
    exp = np.zeros(13 * 11)
    exp[keep] = M.ravel(order="F")[keep]
    exp = exp / np.linalg.norm(exp)
    assert np.allclose(got.ravel(order="F"), exp, rtol=1e-15, atol=1e-16)


@pytest.mark.parametrize("vec,k,expected", [
    # hand traces of heap selection (make_heap over the first k, then |x| > |top| replaces the top)
    ([1, 1, 1, 1], 2, {0, 1}),
    ([1, 1, 1, 1, 2], 2, {0, 4}),
    ([1, 1, 1, 2], 3, {0, 1, 3}),
    ([1, -1, 1, -1, 1, 3], 1, {5}),
    ([2, 2, 2, 2], 1, {0}),
])
def test_prox_ties_hand_traces(vec, k, expected):
    M = np.array(vec, dtype=float).reshape(1, -1)
    got = oracle.prox("splin", M, k, normalized=False)
    assert set(np.nonzero(got[0])[0]) == expected


def test_prox_splincol_union_and_normalisation():
    rng = np.random.default_rng(7)
    M = rng.standard_normal((10, 10))
    c = oracle.prox("spcol", M, 2, normalized=False)
    l = oracle.prox("splin", M, 2, normalized=False)
    u = oracle.prox("splincol", M, 2, normalized=True)
    union = np.where(c != 0, c, l)
    assert np.allclose(u, union / np.linalg.norm(union), rtol=1e-15, atol=1e-16)
    with pytest.raises(RuntimeError):
        oracle.prox("sp", M, 0)  # zero matrix: MatDense::normalize throws


def test_prox_pos_and_supp():
    M = np.array([[1.0, -2.0], [-3.0, 4.0]])
    got = oracle.prox("sp", M, 2, normalized=False, pos=True)
    assert np.array_equal(got, np.array([[1.0, 0.0], [0.0, 4.0]]))
    S = np.array([[1.0, 0.0], [1.0, 1.0]])
    got = oracle.prox("supp", M, supp=S, normalized=False)
    assert np.array_equal(got, M * S)


# ---- PALM4MSA / hierarchical ------------------------------------------------
@pytest.mark.parametrize("nl", [4, 5])
def test_hierarchical_recovers_hadamard_exactly(nl):
    n = 2 ** nl
    H = oracle.hadamard(n)
    fac, res, kw = oracle.wht_params(nl, 30)
    r = oracle.hierarchical(H, fac, res, **kw)
    P = np.eye(n)
    for f in r.factors:
        P = P @ f
    assert np.linalg.norm(r.lam * P - H) / np.linalg.norm(H) < 1e-12
    # each factor keeps 2 nonzeros per row and per column (splincol k=2 on a butterfly)
    for f in r.factors:
        assert (np.count_nonzero(f, axis=1) == 2).all() and (np.count_nonzero(f, axis=0) == 2).all()


def test_palm_lambda_is_the_least_squares_scale():
    # update_lambda: lambda = tr(A^T S)/||S||^2 with S the product of the final
    # factors (left-to-right sweep: the packed left product holds the new S_0).
    # With right-to-left updates the reference multiplies the left product of
    # the start of the sweep instead (faust_palm4msa2020.hpp:450-467), so this
    # identity is only checked for the left-to-right order.
    rng = np.random.default_rng(8)
    A = rng.standard_normal((20, 20))
    r = oracle.palm4msa(A, [("splin", 5, 20, 20), ("spcol", 5, 20, 20)], num_its=12, packing_RL=False,
                        is_update_way_R2L=False)
    S = r.factors[0] @ r.factors[1]
    lam = np.trace(A.T @ S) / np.linalg.norm(S) ** 2
    assert abs(lam - r.lam) < 1e-12 * abs(lam)


def test_palm_objective_decreases_on_average():
    rng = np.random.default_rng(9)
    A = rng.standard_normal((24, 24))
    r = oracle.palm4msa(A, [("sp", 150, 24, 24), ("sp", 150, 24, 24)], num_its=40, record_traj=True)
    e = r.traj[:, 2]
    assert e[-1] < e[0]
    assert np.all(e > 0) and np.all(e < 2)


def test_grad1_agrees_with_grad2():
    # compute_n_apply_grad1 and compute_n_apply_grad2 are two groupings of the
    # same update; the restatement of grad1 must follow grad2 to rounding
    rng = np.random.default_rng(21)
    A = np.asfortranarray(rng.standard_normal((32, 24)))
    cons = [("splin", 4, 32, 16), ("sp", 64, 16, 16), ("spcol", 3, 16, 24)]
    for R2L in (False, True):
        for packing in (True, False):
            kw = dict(num_its=15, record_traj=True, is_update_way_R2L=R2L, packing_RL=packing)
            r2 = oracle.palm4msa(A, cons, **kw)
            r1 = oracle.palm4msa(A, cons, use_grad1=True, **kw)
            assert abs(r1.lam - r2.lam) <= 1e-12 * abs(r2.lam)
            assert np.abs(np.array(r1.traj)[:, 2] - np.array(r2.traj)[:, 2]).max() < 1e-12
            for a, b in zip(r1.factors, r2.factors):
                assert np.abs(a - b).max() < 1e-12


def _trees(i, j):
    """All parenthesizations of factors i..j as (cost-free) nested tuples."""
    if i == j:
        yield i
        return
    for k in range(i, j):
        for left in _trees(i, k):
            for right in _trees(k + 1, j):
                yield (left, right, i, k, j)


def test_dynprog_order_is_the_cheapest_parenthesization():
    # enumerate every binary tree over 2..6 mixed dense/sparse factors; the
    # cost of joining (i..k)(k+1..j) depends on the formats of factors i and j
    # and on the dimensions (faust_prod_opt.cpp:6-86, faust_prod_opt_gen.hpp:81-200)
    rng = np.random.default_rng(8)
    for trial in range(40):
        n = int(rng.integers(2, 7))
        dims = rng.integers(3, 12, size=n + 1)
        facs = []
        for t in range(n):
            a = rng.standard_normal((dims[t], dims[t + 1])) * (rng.random((dims[t], dims[t + 1])) < 0.3)
            facs.append(np.asfortranarray(a) if rng.random() < 0.5 else sp.csr_matrix(a))
        dense = [isinstance(f, np.ndarray) for f in facs]
        nnz = [int(np.count_nonzero(f)) if dense[t] else f.nnz for t, f in enumerate(facs)]

        def join(i, k, j):
            if dense[i] and dense[j]:
                return dims[i] * dims[k + 1] * dims[j + 1]
            if dense[i]:
                return dims[i] * nnz[j]
            if dense[j]:
                return nnz[i] * dims[j + 1]
            return nnz[i] * nnz[j] // dims[i]

        def tree_cost(t):
            if isinstance(t, (int, np.integer)):
                return 0
            left, right, i, k, j = t
            return tree_cost(left) + tree_cost(right) + join(i, k, j)

        best = min(tree_cost(t) for t in _trees(0, n - 1))
        cost, split = oracle.dynprog_tables(facs)
        assert cost[0, n - 1] == best

        def chosen(i, j):
            if i == j:
                return 0
            k = split[i, j]
            return chosen(i, k) + chosen(k + 1, j) + join(i, k, j)

        assert chosen(0, n - 1) == best


def _load_h32_config():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "hadamard32_config.json")) as f:
        return json.load(f)


def test_hierarchical2016_reference_known_answer():
    # the reference's own acceptance test of its 2016 hierarchical
    # factorization on the 32x32 Hadamard matrix (config_HADAMARD_32.mat;
    # lambda = 5930 +- 30, misc/test/CMakeLists.txt:339), on the oracle's
    # restatement of Faust::HierarchicalFact / Faust::Palm4MSA
    cfg = _load_h32_config()
    H = oracle.hadamard(cfg["nrow"])
    fac = [tuple(c) for c in cfg["fact_constraints"]]
    res = [tuple(c) for c in cfg["res_constraints"]]
    r = oracle.hierarchical2016(H, fac, res, num_its_local=cfg["niter1"], num_its_global=cfg["niter2"],
                                is_update_way_R2L=bool(cfg["update_way"]))
    assert len(r.factors) == cfg["nfacts"]
    assert abs(r.lam - cfg["expected_lambda"]) <= cfg["lambda_tolerance"], r.lam
    P = np.eye(cfg["nrow"])
    for f in r.factors:
        P = P @ f
    assert np.linalg.norm(r.lam * P - H) / np.linalg.norm(H) < 1e-10
    assert all(int((f != 0).sum()) == 64 for f in r.factors)  # splincol 2 on 32x32


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("R2L", [False, True])
def test_palm4msa2016_lambda_closed_form(mode, R2L):
    # compute_lambda's closed form on the returned factors:
    # lambda = |tr(A^T P)| / tr(P^T P), P the product of the final factors
    rng = np.random.default_rng(7)
    A = rng.standard_normal((24, 20))
    cons = [("splin", 4, 24, 16), ("spcol", 3, 16, 16), ("sp", 60, 16, 20)]
    r = oracle.palm4msa2016(A, cons, num_its=12, record_traj=True, is_update_way_R2L=R2L, grad_calc_opt_mode=mode)
    P = r.factors[0] @ r.factors[1] @ r.factors[2]
    lam = abs(np.trace(A.T @ P)) / np.trace(P.T @ P)
    assert abs(r.lam - lam) <= 1e-12 * lam
    assert r.traj.shape == (12, 4) and r.traj[-1, 3] == r.lam
    # the projectors' supports: splin 4 per row, spcol 3 per column, sp 60
    assert np.all((r.factors[0] != 0).sum(axis=1) <= 4)
    assert np.all((r.factors[1] != 0).sum(axis=0) <= 3)
    assert (r.factors[2] != 0).sum() <= 60
    # the iterations decrease the error against A (a descent method)
    err = np.linalg.norm(r.lam * P - A) / np.linalg.norm(A)
    r1 = oracle.palm4msa2016(A, cons, num_its=1, is_update_way_R2L=R2L, grad_calc_opt_mode=mode)
    P1 = r1.factors[0] @ r1.factors[1] @ r1.factors[2]
    assert err < np.linalg.norm(r1.lam * P1 - A) / np.linalg.norm(A)


def test_palm4msa2016_gradient_modes_agree():
    # the three gradient forms associate the same product differently: the
    # same iterates up to rounding
    rng = np.random.default_rng(8)
    A = rng.standard_normal((16, 16))
    cons = [("splincol", 3, 16, 16), ("splincol", 3, 16, 16)]
    rs = [oracle.palm4msa2016(A, cons, num_its=3, grad_calc_opt_mode=m) for m in (0, 1, 2)]
    for r in rs[1:]:
        assert abs(r.lam - rs[0].lam) <= 1e-9 * rs[0].lam
        for a, b in zip(r.factors, rs[0].factors):
            assert np.allclose(a, b, rtol=0, atol=1e-9 * np.abs(b).max())
