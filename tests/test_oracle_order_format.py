"""Pins of two rules the oracle restates from the reference (CPU).

* multiply_order_opt_all_ends (faust_prod_opt.hpp:49-86), the product order
  of PALM4MSA's gradient: while more than two factors are left, compare
  R1.rows * R1.cols * R2.cols (the two left-most) with L1.rows * L1.cols *
  L2.cols (the two right-most), stored dimensions whatever the flags, and
  multiply the left pair when the first is smaller, else the right pair.
  The factors here are generalized permutation matrices with values in
  [1, 2): every entry of a product is one product of scalars, so numpy
  computes each association exactly as the oracle's canonical order does,
  and different associations round differently.  Every end-pair sequence is
  enumerated; the oracle must give the bits of the one the rule picks, and
  the cases are chosen so that other sequences give other bits.
* the factor format of project_gen (faust_prox_gen.hpp): dense when
  sparse_size(nnz, rows) > dense_size(rows, cols) with the byte counts of
  faust_prox.hpp:98-110 and, for the sparsity projectors, the nnz of the
  constraint's parameter (k for sp, k * rows for spcol / splincol, k * cols
  for splin), the real nonzeros for the others.  The reference's own case
  (misc/test/src/C++/faust_sparse_prox_sp.cpp.in:28-51: 204 x 204, k = 1632,
  the result taken as a MatSparse and equal to prox_sp) and the boundaries of
  each rule.
"""
import itertools

import numpy as np
import pytest

import oracle


def _gperm(m, n, rng):
    """m x n, at most one nonzero per row and column, values in [1, 2)."""
    A = np.zeros((m, n))
    k = min(m, n)
    rows = rng.permutation(m)[:k]
    cols = rng.permutation(n)[:k]
    A[rows, cols] = 1.0 + rng.random(k)
    return A


def _op(A, f):
    return A.T if f in "TH" else A


def _end_pair_result(facs, flags, seq):
    """The product when step s multiplies the left-most pair (seq[s] = 0) or
    the right-most pair (1); intermediate products are plain (flag N)."""
    items = [(_op(F, flags[i] if len(flags) > i else flags[0]), F.shape) for i, F in enumerate(facs)]
    for side in seq:
        if side == 0:
            (a, _), (b, _) = items[0], items[1]
            p = a @ b
            items = [(p, p.shape)] + items[2:]
        else:
            (a, _), (b, _) = items[-2], items[-1]
            p = a @ b
            items = items[:-2] + [(p, p.shape)]
    (a, _), (b, _) = items
    return a @ b


CASES = [
    # (stored shapes, flags, what the case checks); every product has at least
    # 8 nonzero entries, so that associations are told apart by their bits
    (((16, 64), (64, 64), (64, 128)), "N", "left pair first (65536 < 524288)"),
    (((128, 64), (64, 64), (64, 16)), "N", "right pair first"),
    (((32, 32), (32, 32), (32, 32)), "N", "a tie goes to the right pair"),
    (((32, 16), (32, 32), (8, 32)), "TNT", "stored shapes: L^T (16x32) res (32x32) R^T (32x8); the stored counts "
                                          "32*16*32 < 32*32*32 take the left pair, the op'd ones (16384 > 8192) "
                                          "would take the right"),
    (((8, 12), (12, 20), (20, 28), (28, 44)), "N", "two steps on the left"),
    (((44, 28), (28, 20), (20, 12), (12, 8)), "N", "two steps on the right"),
    (((24, 8), (8, 36), (36, 12), (12, 32), (32, 20)), "N", "left, right, right"),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_mul_all_ends_order_pinned_by_bits(case):
    shapes, flags, _ = CASES[case]
    rng = np.random.default_rng(100 + case)
    facs = [_gperm(m, n, rng) for m, n in shapes]
    fl = flags if len(flags) == len(facs) else "N" * len(facs)
    opd = [_op(F, fl[i]).shape for i, F in enumerate(facs)]
    # the shapes the rule compares: stored shapes of the factors, and of the
    # running products (stored as computed, i.e. op'd)
    shp = [F.shape for F in facs]
    ri, li = 0, len(facs) - 1
    seq = []
    while li - ri > 1:
        (r1m, r1n), (_, r2n) = shp[ri], shp[ri + 1]
        (l1m, l1n), (_, l2n) = shp[li - 1], shp[li]
        if r1m * r1n * r2n < l1m * l1n * l2n:
            seq.append(0)
            shp[ri + 1] = (opd[0][0], opd[ri + 1][1])
            ri += 1
        else:
            seq.append(1)
            shp[li - 1] = (opd[li - 1][0], opd[-1][1] if li == len(facs) - 1 else shp[li][1])
            li -= 1
    want = _end_pair_result(facs, fl, seq)
    got = oracle.mul_all_ends(facs, flags=fl if flags != "N" else "N")
    assert np.array_equal(got, want), (case, seq)
    # the case discriminates: some other end-pair sequence rounds differently
    others = [s for s in itertools.product((0, 1), repeat=len(facs) - 2) if list(s) != seq]
    assert any(not np.array_equal(_end_pair_result(facs, fl, list(s)), want) for s in others), case


def test_mul_all_ends_alpha_beta():
    rng = np.random.default_rng(7)
    facs = [_gperm(6, 6, rng) for _ in range(3)]
    C0 = rng.standard_normal((6, 6))
    got = oracle.mul_all_ends(facs, "N", alpha=0.5, beta=-1.0, out=C0)
    P = oracle.mul_all_ends(facs, "N")
    assert np.allclose(got, 0.5 * P - C0, rtol=0, atol=1e-15)


def _sizes(nnz, rows, cols):
    # faust_prox.hpp:98-110 at FPP = double, int = 4 bytes
    return nnz * (8 + 4) + (rows + 1) * 4, rows * cols * 8


def test_project_gen_format_reference_case():
    # faust_sparse_prox_sp.cpp.in: 204 x 204, values randMat * 100, k = 1632:
    # prox_sp_gen's result is used as a MatSparse equal to prox_sp's
    rng = np.random.default_rng(0)
    M = rng.random((204, 204)) * 100
    assert oracle.project_gen_kind("sp", M, 1632, normalized=False) == "sparse"
    assert np.array_equal(oracle.prox("sp", M, 1632, normalized=False),
                          oracle.prox("sp", M, 1632, normalized=False))


@pytest.mark.parametrize("name,shape,nnz_of_k", [
    ("sp", (204, 204), lambda k, m, n: k),
    ("spcol", (40, 10), lambda k, m, n: k * m),     # k * dim1 (faust_prox_gen.hpp:137)
    ("splincol", (40, 10), lambda k, m, n: k * m),  # k * dim1 (:123)
    ("splin", (10, 40), lambda k, m, n: k * n),     # k * dim2 (:150)
])
def test_project_gen_format_boundary(name, shape, nnz_of_k):
    m, n = shape
    rng = np.random.default_rng(1)
    M = rng.standard_normal(shape)
    kinds = {}
    limit = m * n if name == "sp" else (m if name == "spcol" else n)
    for k in range(1, limit + 1):
        s, d = _sizes(nnz_of_k(k, m, n), m, n)
        kinds[k] = "dense" if s > d else "sparse"
    assert "dense" in kinds.values() and "sparse" in kinds.values()  # the boundary lies inside the range
    edge = max(k for k, v in kinds.items() if v == "sparse")
    for k in sorted({1, edge, edge + 1, limit} & set(kinds)):
        assert oracle.project_gen_kind(name, M, k) == kinds[k], (name, k)


def test_project_gen_format_real_nonzeros():
    # supp / normcol: the nonzeros left after the projection decide
    rng = np.random.default_rng(2)
    M = rng.standard_normal((20, 20))
    S = np.zeros((20, 20))
    S[:, :3] = 1  # 60 nonzeros: 12 * 60 + 84 = 804 <= 3200 -> sparse
    assert oracle.project_gen_kind("supp", M, supp=S) == "sparse"
    S[:, :17] = 1  # 340: 4164 > 3200 -> dense
    assert oracle.project_gen_kind("supp", M, supp=S) == "dense"
    assert oracle.project_gen_kind("normcol", M, 1.0) == "dense"
    Z = M.copy()
    Z[:, 2:] = 0  # 40 nonzeros left after normcol -> sparse
    assert oracle.project_gen_kind("normcol", Z, 1.0) == "sparse"
