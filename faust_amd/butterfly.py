"""Butterfly supports, random butterfly Fausts and the permutations that
pyfaust's butterfly factorization takes (pyfaust.rand_butterfly,
pyfaust.py:4810-4848; pyfaust.bitrev_perm, pyfaust.py:3806-3824; the
'default_8' permutations of pyfaust.fact.butterfly, fact.py:1813-1873).

Support i of L = log2(n) is kron(I_{2^i}, kron(ones(2, 2), I_{2^(L-i-1)}))
(faust_butterfly.hpp:459-490): row r holds columns
c0 = (r & ~(2s - 1)) + (r & (s - 1)) and c0 + s, with s = 2^(L-i-1).

A permutation is handled as pyfaust handles it at the C++ boundary: an index
vector perm with P[perm[j], j] = 1 (perm[j] is the row of the one in column
j; perm2indices(P) = P.T.nonzero()[1]).
"""
import numpy as np
import scipy.sparse as sp


def _log2(n):
    L = int(n).bit_length() - 1
    if n < 1 or (1 << L) != n:
        raise ValueError("n must be a power of two")
    return L


def support_columns(n, i):
    """(c0, c1): the two columns row r of butterfly support i of order n holds."""
    L = _log2(n)
    s = 1 << (L - i - 1)
    r = np.arange(n)
    c0 = (r & ~(2 * s - 1)) + (r & (s - 1))
    return c0, c0 + s


def support_csr(n, i, values=None):
    """Butterfly support i of order n as a CSR matrix (values: 2n entries in
    row order, ones when None)."""
    c0, c1 = support_columns(n, i)
    ci = np.stack([c0, c1], axis=1).reshape(-1)
    v = np.ones(2 * n) if values is None else np.asarray(values).reshape(-1)
    return sp.csr_matrix((v, ci, np.arange(0, 2 * n + 1, 2)), shape=(n, n))


def bitrev(n):
    """Bit-reversal of 0..n-1 on log2(n) bits."""
    L = _log2(n)
    i = np.arange(n)
    r = np.zeros(n, dtype=np.int64)
    for b in range(L):
        r |= ((i >> b) & 1) << (L - 1 - b)
    return r


def bitrev_perm(n):
    """pyfaust.bitrev_perm: the CSR permutation with P[i, bitrev(i)] = 1."""
    return sp.csr_matrix((np.ones(n), (np.arange(n), bitrev(n))), shape=(n, n))


def perm2indices(P):
    """Index form of a permutation matrix: the row of the one in each column."""
    return sp.csr_matrix(P).T.nonzero()[1]


def indices2perm(perm):
    n = len(perm)
    return sp.csr_matrix((np.ones(n), (np.asarray(perm), np.arange(n))), shape=(n, n))


# Index form r of a permutation matrix Q (Q e_j = e_{r[j]}); a product Q1 Q2
# is r1[r2], and kron(I_m, Q) acts blockwise.
def _type_c(size):
    """First half fixed, second half reversed (columns)."""
    h = size // 2
    r = np.arange(size)
    r[h:] = h + (size - 1 - np.arange(h, size))
    return r


def _type_b(size):
    """First half reversed (rows), second half fixed."""
    h = size // 2
    r = np.arange(size)
    r[:h] = h - 1 - np.arange(h)
    return r


def _type_a(size):
    """Even columns to the first half, odd columns to the second."""
    h = size // 2
    r = np.empty(size, dtype=np.int64)
    r[0::2] = np.arange(h)
    r[1::2] = h + np.arange(h)
    return r


def _blockwise(n, r_block):
    b = len(r_block)
    j = np.arange(n)
    return (j // b) * b + r_block[j % b]


def default_8_perms(n):
    """The eight permutations pyfaust.fact.butterfly tries for
    perm='default_8': for each choice c in 000..111, the product over levels
    i = 2..L of kron(I_{2^(L-i)}, Q_i), Q_i the product of the level's type c,
    b, a permutations selected by the three digits (fact.py:1813-1873)."""
    L = _log2(n)
    out = []
    for name in ["000", "001", "010", "011", "100", "101", "110", "111"]:
        ch = [int(x) for x in name]
        total = np.arange(n)
        for i in range(2, L + 1):
            size = 1 << i
            blk = np.arange(size)
            for use, mk in zip(ch, (_type_c, _type_b, _type_a)):
                if use:
                    blk = blk[mk(size)]
            total = total[_blockwise(n, blk)]
        out.append(total)
    return out


def rand_butterfly_factors(n, dtype=np.float64, rng=None):
    """The log2(n) butterfly supports of order n with standard normal values
    (np.random.randn, or rng.standard_normal), as CSR matrices."""
    L = _log2(n)
    draw = (lambda k: np.random.randn(k)) if rng is None else rng.standard_normal
    return [support_csr(n, i, draw(2 * n)).astype(dtype) for i in range(L)]
