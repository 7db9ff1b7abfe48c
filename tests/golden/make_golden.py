"""Writes the golden fixtures in this folder from exact rational arithmetic.

The reference (DDLA, C#) ships no golden vectors and cannot run here, so every fixture is a
known-answer case whose expected output follows from the mathematics alone:

  gemm_int.npz      small-integer A (7x9... up to 300 deep), B: every partial sum is an
                    exactly representable integer, so C = A B holds bit for bit in FP64
  gepp_exact.npz    partial-pivoting LU computed in exact fractions (first index of max
                    |x| wins, relative pivots as in DDLA/Factorizations/LU.cs:367-374) for
                    small random matrices whose pivot choices are not near ties
  lu_designed.npz   A = P^-1 L U from integer unit-lower L (|l| <= 1 so partial pivoting
                    keeps the designed order) and integer U with power-of-two diagonal:
                    P A = L U is exact in FP64 and the pivots are known
  chol_int.npz      A = L L^T from integer L with power-of-two diagonal (sqrt and the
                    scaling are exact), lower factor known exactly

Run: python tests/golden/make_golden.py   (numpy + fractions only)
"""
import os
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def gepp_exact(A):
    """Exact GEPP; returns (packed LU as Fractions, relative pivots, min pivot gap ratio)."""
    n = len(A)
    M = [[Fraction(x) for x in row] for row in A]
    piv = []
    gap = float("inf")
    for i in range(n):
        col = [abs(M[r][i]) for r in range(i, n)]
        best = max(col)
        p = col.index(best)  # first index of the maximum
        others = sorted(set(col), reverse=True)
        if len(others) > 1 and best != 0:
            gap = min(gap, float((others[0] - others[1]) / best))
        piv.append(p)
        if p:
            M[i], M[i + p] = M[i + p], M[i]
        if M[i][i] == 0:
            raise ZeroDivisionError
        for r in range(i + 1, n):
            M[r][i] /= M[i][i]
            for c in range(i + 1, n):
                M[r][c] -= M[r][i] * M[i][c]
    return M, np.array(piv, dtype=np.int32), gap


def main():
    rng = np.random.default_rng(20261203)
    # GEMM known answers
    shapes = [(7, 9, 300), (13, 17, 145), (6, 8, 144), (1, 1, 1), (33, 5, 289)]
    g = {}
    for t, (m, n, k) in enumerate(shapes):
        A = rng.integers(-8, 9, size=(m, k)).astype(np.float64)
        B = rng.integers(-8, 9, size=(k, n)).astype(np.float64)
        C = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
        g[f"A{t}"], g[f"B{t}"], g[f"C{t}"] = A, B, C
    g["count"] = np.array(len(shapes))
    np.savez_compressed(os.path.join(HERE, "gemm_int.npz"), **g)

    # exact GEPP on small random matrices (dyadic entries so the input is exact)
    e = {}
    cnt = 0
    for n in (3, 5, 8, 12):
        for trial in range(3):
            A = rng.integers(-1024, 1025, size=(n, n)) / 1024.0
            try:
                M, piv, gap = gepp_exact(A)
            except ZeroDivisionError:
                continue
            if gap < 1e-6:
                continue  # near tie: floating point may legitimately choose differently
            e[f"A{cnt}"] = A
            e[f"piv{cnt}"] = piv
            e[f"LU{cnt}"] = np.array([[float(x) for x in row] for row in M])
            cnt += 1
    e["count"] = np.array(cnt)
    np.savez_compressed(os.path.join(HERE, "gepp_exact.npz"), **e)

    # designed LU: P A = L U exact, pivots known
    d = {}
    for t, n in enumerate((6, 130, 257)):
        L = np.tril(rng.integers(-1, 2, size=(n, n)).astype(np.float64), -1) * 0.5
        np.fill_diagonal(L, 1.0)
        U = np.triu(rng.integers(-3, 4, size=(n, n)).astype(np.float64), 1)
        np.fill_diagonal(U, 2.0 ** rng.integers(1, 4, size=n))
        # with |l| <= 1/2 the largest entry of each column below the diagonal is the
        # diagonal row itself, so with P = I the pivots are 0; build a permuted A whose
        # pivot sequence is the designed one
        perm_rel = rng.integers(0, 3, size=n)
        perm_rel = np.minimum(perm_rel, n - 1 - np.arange(n)).astype(np.int32)
        LUp = L @ U  # exact: small integers and halves
        A = LUp.copy()
        # apply the inverse permutation (backward interchanges) so P A = L U
        for i in range(n - 1, -1, -1):
            p = perm_rel[i]
            if p > 0:
                A[[i, i + p]] = A[[i + p, i]]
        d[f"A{t}"], d[f"L{t}"], d[f"U{t}"], d[f"piv{t}"] = A, L, U, perm_rel
    d["count"] = np.array(3)
    np.savez_compressed(os.path.join(HERE, "lu_designed.npz"), **d)

    # Cholesky known answers
    c = {}
    for t, n in enumerate((4, 129, 300)):
        L = np.tril(rng.integers(-2, 3, size=(n, n)).astype(np.float64), -1)
        np.fill_diagonal(L, 2.0 ** rng.integers(2, 5, size=n))
        A = L @ L.T
        c[f"A{t}"], c[f"L{t}"] = A, L
    c["count"] = np.array(3)
    np.savez_compressed(os.path.join(HERE, "chol_int.npz"), **c)
    print("fixtures written:", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
