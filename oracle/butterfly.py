"""FP64 CPU restatement of the butterfly factorization (TEST INFRASTRUCTURE ONLY).

Imported by tests/ as the checker; the product package never imports it.

Follows src/algorithm/factorization/faust_butterfly.hpp:
  * butterfly_support (:459-490): support i of L is
    kron(I_{2^i}, kron(ones(2, 2), I_{2^(L-i-1)}));
  * lifting_two_layers_factorization (:176-246): for every column t of the
    left support, the rows t touches there and the columns t touches in the
    right support select a block of A; its best rank-1 approximation
    u*sqrt(s) / sqrt(s)*v^T (MatDense::best_low_rank) becomes column t of X
    and row t of Y;
  * retrieveCEC / solveDTO (:44-96, :336-455): the directional driver's step;
  * butterfly_hierarchical (:545-595): 'right' / 'left' chains of solveDTO;
  * butterfly_hierarchical_balanced (:598-700): the tree of supports;
  * butterfly_hierarchical(A, dir, P, mul_perm) (:804-860) and
    pyfaust.fact.butterfly (wrapper/python/pyfaust/fact.py:1793-1907): the
    permutation forms.
Supports and blocks are found here the way the reference finds them: dense
0/1 masks, products of masks, and the nonzero indices of their rows and
columns.

Best rank-1 approximations come from numpy's LAPACK SVD.  Singular vectors
are defined up to sign, and Eigen (the reference) and LAPACK may pick
different ones: factors agree up to a +-1 diagonal between consecutive
factors, so comparisons use |factor| and products.
"""
import numpy as np
import scipy.sparse as sp

LEFT, RIGHT, BALANCED = 0, 1, 2  # ButterflyFactDir, faust_butterfly.h:5-10


def butterfly_supports(L):
    """The L butterfly supports of order 2^L, dense 0/1 (faust_butterfly.hpp:459-490)."""
    out = []
    for i in range(L):
        kernel = np.kron(np.ones((2, 2)), np.eye(1 << (L - i - 1)))
        out.append(np.kron(np.eye(1 << i), kernel))
    return out


def _nz1(M):
    """setNZtoOne: a support mask from a product of masks."""
    return (M != 0).astype(np.float64)


def best_rank1(M):
    """(x, y) with x y^T the best rank-1 approximation of M:
    x = u_1 sqrt(s_1), y = sqrt(s_1) v_1 (MatDense::best_low_rank, rank 1)."""
    U, s, Vt = np.linalg.svd(M, full_matrices=False)
    r = np.sqrt(s[0])
    return U[:, 0] * r, Vt[0, :] * r


def best_low_rank(M, k):
    U, s, Vt = np.linalg.svd(M, full_matrices=False)
    r = np.sqrt(s[:k])
    return U[:, :k] * r, (Vt[:k, :].T * r).T


def lift(A, S1, S2):
    """lifting_two_layers_factorization: A ~ X Y with supp(X) in S1, supp(Y) in S2."""
    r = S1.shape[1]
    X = np.zeros((S1.shape[0], r))
    Y = np.zeros((r, S2.shape[1]))
    for t in range(r):
        rows = np.nonzero(S1[:, t])[0]
        cols = np.nonzero(S2[t, :])[0]
        x, y = best_rank1(A[np.ix_(rows, cols)])
        X[rows, t] = x
        Y[t, cols] = y
    return X, Y


def cec_classes(S1, S2):
    """retrieveCEC: indices t grouped when column t of S1 and row t of S2 both
    match; a class goes to 'cec' when min(|row sum|, |column sum|) of S1 at t
    is at most its size.  Returns (cec, noncec) lists of index lists."""
    # supports are 0/1, so the reference's comparisons within 1e-6 are
    # equality of patterns: group equal (column of S1, row of S2) pairs,
    # classes in order of their first index, members ascending
    pat = np.concatenate([S1.T, S2], axis=1)
    _, first, inv = np.unique(pat, axis=0, return_index=True, return_inverse=True)
    inv = np.asarray(inv).reshape(-1)
    cec, noncec = [], []
    for g in np.argsort(first, kind="stable"):
        cls = [int(j) for j in np.nonzero(inv == g)[0]]
        i = cls[0]
        if min(abs(S1[i, :].sum()), abs(S1[:, i].sum())) <= len(cls):
            cec.append(cls)
        else:
            noncec.append(cls)
    return cec, noncec


def solve_dto(A, S1, S2):
    """solveDTO for supports whose classes are all non-CEC (every butterfly
    pair of supports): each class of size k takes the best rank-k
    approximation of its block."""
    cec, noncec = cec_classes(S1, S2)
    if cec:
        raise NotImplementedError("solveDTO: complete equivalence classes do not arise with butterfly supports")
    X = np.zeros((S1.shape[0], S1.shape[1]))
    Y = np.zeros((S2.shape[0], S2.shape[1]))
    for cls in noncec:
        r = cls[0]
        rows = np.nonzero(S1[:, r])[0]
        cols = np.nonzero(S2[r, :])[0]
        bx, by = best_low_rank(A[np.ix_(rows, cols)], len(cls))
        for i, c in enumerate(cls):
            X[rows, c] = bx[:, i]
            Y[c, cols] = by[i, :]
    return X, Y


def hierarchical_dir(A, supports, direction):
    """butterfly_hierarchical for 'right' (peel factors off the left end) or
    'left' (off the right end).  Returns dense factors in product order."""
    n_s = len(supports)
    s2 = np.eye(A.shape[0])
    s2_vec = [None] * n_s
    i = 0 if direction == RIGHT else n_s - 1
    j = n_s - 1 if direction == RIGHT else 0
    while True:
        s2 = _nz1(supports[j] @ s2)
        s2_vec[j - 1 if direction == RIGHT else j + 1] = s2
        if direction == RIGHT:
            j = j - 1 if j > i + 1 else -1
        else:
            j = j + 1 if j < i - 1 else -1
        if j < 0:
            break
    mat = A
    facs = []
    while True:
        s1 = supports[i]
        if direction == RIGHT:
            X, Y = solve_dto(mat, s1, s2_vec[i])
            facs.append(X)
        else:
            Y, X = solve_dto(mat, s2_vec[i], s1)
            facs.insert(0, X)
        mat = Y
        if direction == RIGHT:
            i = i + 1 if i < n_s - 2 else -1
        else:
            i = i - 1 if i > 1 else -1
        if i < 0:
            break
    if direction == RIGHT:
        facs.append(mat)
    else:
        facs.insert(0, mat)
    return facs


def tree_supports(supports):
    """The support tree of butterfly_hierarchical_balanced: level 0 has two
    nodes, each inner level doubles, the last level holds the remaining
    leaves; a node is a leaf support or the product of its two children."""
    n = len(supports)
    d = int(np.ceil(np.log2(n)))
    sizes = [1 << (i + 1) for i in range(d - 1)]
    sizes.append(2 * (n - sizes[d - 2]) if d > 1 else 2)
    tree = [[None] * s for s in sizes]
    for i in range(d - 1, -1, -1):
        for j in range(sizes[i]):
            if i == d - 1:
                tree[i][j] = supports[j]
            elif 2 * (j + 1) > sizes[i + 1]:
                tree[i][j] = supports[j + sizes[i + 1] // 2]
            else:
                tree[i][j] = _nz1(tree[i + 1][2 * j] @ tree[i + 1][2 * j + 1])
    return tree


def balanced(A, supports):
    """butterfly_hierarchical_balanced (lifting at every node).  Returns
    dense factors in product order."""
    tree = tree_supports(supports)
    d = len(tree)
    nxt = list(lift(A, tree[0][0], tree[0][1]))
    inp = nxt
    for i in range(d - 1):
        nxt = [None] * len(tree[i + 1])
        for j in range(min(len(tree[i]), len(tree[i + 1]) // 2)):
            nxt[2 * j], nxt[2 * j + 1] = lift(inp[j], tree[i + 1][2 * j], tree[i + 1][2 * j + 1])
        if i < d - 2:
            inp = nxt
    facs = list(nxt)
    if len(nxt) < len(supports):
        facs += inp[len(nxt) // 2:]
    return facs


def perm_matrix(perm):
    """P with P[perm[i], i] = 1 (FaustFact.hpp:905-918)."""
    n = len(perm)
    return sp.csr_matrix((np.ones(n), (np.asarray(perm), np.arange(n))), shape=(n, n))


def butterfly(M, direction=BALANCED, perm=None, mul_perm=True):
    """butterfly_hierarchical(A, dir, P, mul_perm): factorize A P^T, then put P
    back on the right (multiplied into the last factor, or as a factor of its
    own).  Returns the factors in product order (dense arrays; a permutation
    factor or a permuted last factor as scipy CSR)."""
    A = np.asarray(M, dtype=np.float64)
    n = A.shape[0]
    L = int(round(np.log2(n)))
    if A.shape[0] != A.shape[1] or (1 << L) != n or n < 4:
        raise ValueError("M must be square of order a power of two >= 4")
    sup = butterfly_supports(L)
    P = None
    if perm is not None:
        P = perm_matrix(perm)
        A = np.asarray(A @ P.T)
    facs = balanced(A, sup) if direction == BALANCED else hierarchical_dir(A, sup, direction)
    if P is not None:
        if mul_perm:
            facs[-1] = sp.csr_matrix(facs[-1]) @ P
        else:
            facs.append(P)
    return facs


def product(facs):
    out = None
    for f in facs:
        f = f.toarray() if sp.issparse(f) else f
        out = f if out is None else out @ f
    return out


def rel_err(facs, M):
    return np.linalg.norm(product(facs) - M) / np.linalg.norm(M)
