"""Pins the CPU oracle (oracle/ddla_oracle.cpp) before it is trusted as the parity checker.

Each check is chosen so that a plausible slip in the restatement (a dropped term, a wrong
index, a transposed operand, a wrong tie rule or pivot encoding) fails it:
  * exact-arithmetic models (fractions) of the managed GEMM's rounding and of GEPP;
  * known-answer fixtures in tests/golden (written from exact arithmetic);
  * library routines: numpy matmul, LAPACK getrf/potrf (scipy), triangular solves;
  * invariants: P A = L U, A = L L^T, round trips of the interchanges.
"""
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg as sla

from conftest import golden, rel_fro

O = None


@pytest.fixture(autouse=True)
def _o(oracle):
    global O
    O = oracle


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def managed_dot(a, b, kc=144):
    """The managed GEMM's rounding of one element (GEMMKernel.cs:36-240, GeMM.cs:47-49,
    MacroKernel.cs:94): fma chain from 0.0 inside each KC block, block sums added in order."""
    total = 0.0
    for p0 in range(0, len(a), kc):
        s = 0.0
        for k in range(p0, min(p0 + kc, len(a))):
            s = fma(a[k], b[k], s)
        total = total + s
    return total


# ------------------------------------------------------------------------------ GEMM ---
@pytest.mark.parametrize("m,n,k", [(7, 9, 300), (13, 17, 145), (6, 8, 289), (1, 1, 1)])
def test_gemm_bit_exact_vs_fma_chain_model(m, n, k):
    A = O.uniform_matrix(m, k, 11 + m)
    B = O.uniform_matrix(k, n, 17 + n)
    C = O.matmul(A, B)
    for i in range(m):
        for j in range(n):
            assert C[i, j] == managed_dot(A[i], B[:, j]), (i, j)


def test_gemm_fma_chain_model_detects_wrong_blocking():
    # the model with another KC must disagree somewhere: the check above is sensitive to it
    A = O.uniform_matrix(4, 300, 5)
    B = O.uniform_matrix(300, 4, 6)
    C = O.matmul(A, B)
    diffs = sum(C[i, j] != managed_dot(A[i], B[:, j], kc=128) for i in range(4) for j in range(4))
    assert diffs > 0


def test_gemm_integer_known_answers():
    g = golden("gemm_int.npz")
    for t in range(int(g["count"])):
        C = O.matmul(g[f"A{t}"], g[f"B{t}"])
        np.testing.assert_array_equal(C, g[f"C{t}"])


@pytest.mark.parametrize("ta,tb", [(0, 0), (8, 0), (0, 8), (8, 8)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (2.0, 1.2), (-1.0, 1.0), (0.5, -0.75)])
def test_gemm_vs_numpy(ta, tb, alpha, beta):
    m, n, k = 71, 53, 160
    A = O.uniform_matrix(*(k, m) if ta else (m, k), 1)
    B = O.uniform_matrix(*(n, k) if tb else (k, n), 2)
    C0 = O.uniform_matrix(m, n, 3)
    C = C0.copy()
    O.gemm(alpha, A, B, beta, C, ta, tb)
    ref = alpha * ((A.T if ta else A) @ (B.T if tb else B)) + beta * C0
    assert rel_fro(C, ref) < 1e-14 * k


def test_gemm_strided_and_colmajor_views():
    # the reference tests row-major, column-major and colStride=4 views (TestGeMM.cs:13-38)
    m, n, k = 40, 33, 150
    big = O.uniform(4 * m * k + 16, 9)
    A = np.lib.stride_tricks.as_strided(big, shape=(m, k), strides=(8 * 4 * k, 8 * 4))
    B = np.asfortranarray(O.uniform_matrix(k, n, 10))
    C = np.asfortranarray(O.uniform_matrix(m, n, 11))
    ref = 2.0 * (A @ B) + 1.2 * C
    O.gemm(2.0, A, B, 1.2, C)
    assert rel_fro(C, ref) < 1e-14 * k


def test_gemm_beta_zero_clears_nan_and_k_zero_scales():
    A = O.uniform_matrix(5, 0, 1)
    B = O.uniform_matrix(0, 6, 2)
    C = np.full((5, 6), np.nan)
    O.gemm(1.0, A, B, 0.0, C)  # Managed/1m.cs:265-269: beta == 0 -> Set(0)
    assert np.all(C == 0.0)
    C = np.ones((5, 6))
    O.gemm(1.0, A, B, 3.0, C)
    assert np.all(C == 3.0)


# ------------------------------------------------------------------------------ AMAX ---
def test_amax_first_index_and_nan():
    assert O.amax([1.0, -3.0, 3.0, 2.0]) == 1  # first index of the max wins (TestAMaxV.cs)
    assert O.amax([np.nan, 0.5, -0.5]) == 1  # NaN is smaller than every number (1v.cs:30-34)
    assert O.amax([np.nan, np.nan]) == 0
    assert O.amax(np.zeros(0)) == 0


# -------------------------------------------------------------------------------- LU ---
def test_getrf_matches_exact_gepp_fixtures():
    e = golden("gepp_exact.npz")
    assert int(e["count"]) >= 8
    for t in range(int(e["count"])):
        A = e[f"A{t}"].copy()
        info, piv = O.getrf(A)
        assert info == 0
        np.testing.assert_array_equal(piv, e[f"piv{t}"])
        np.testing.assert_allclose(A, e[f"LU{t}"], rtol=0, atol=1e-12)


def test_getrf_designed_known_answers():
    d = golden("lu_designed.npz")
    for t in range(int(d["count"])):
        A = d[f"A{t}"].copy()
        info, piv = O.getrf(A)
        assert info == 0
        np.testing.assert_array_equal(piv, d[f"piv{t}"])
        L, U = O.lu_factors(A)
        np.testing.assert_array_equal(L, d[f"L{t}"])
        np.testing.assert_array_equal(U, d[f"U{t}"])


@pytest.mark.parametrize("n", [5, 127, 128, 130, 257, 600])
def test_getrf_vs_lapack(n):
    A0 = O.uniform_matrix(n, n, 1000 + n)
    A = A0.copy()
    info, piv = O.getrf(A)
    assert info == 0
    lu, ipiv = sla.lu_factor(A0)
    np.testing.assert_array_equal(piv, ipiv - np.arange(n))  # absolute -> relative
    assert np.abs(A - lu).max() < 1e-10
    L, U = O.lu_factors(A)
    P = O.perm_matrix(piv)
    assert rel_fro(P @ A0, L @ U) < 1e-14 * n


def test_getrf_zero_pivot_info():
    A = O.uniform_matrix(6, 6, 3)
    A[:, 2] = A[:, 0] * 2.0  # singular: third column dependent on the first... exactly
    A[:, 2] = 0.0
    info, _ = O.getrf(A)
    assert info == 3


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("n,nrhs", [(200, 1), (300, 7)])
def test_getrs_vs_numpy(n, nrhs, trans):
    A0 = O.uniform_matrix(n, n, 7 + n)
    B = O.uniform_matrix(n, nrhs, 8)
    A = A0.copy()
    _, piv = O.getrf(A)
    X = O.getrs(A, piv, B.copy(), trans=trans)
    ref = np.linalg.solve(A0.T if trans else A0, B)
    assert rel_fro(X, ref) < 1e-10


def test_laswp_roundtrip_and_vs_permutation():
    n = 50
    piv = np.array([np.random.default_rng(1).integers(0, n - i) for i in range(n)], dtype=np.int32)
    X0 = O.uniform_matrix(n, 3, 4)
    X = O.laswp(X0.copy(), piv)
    # same as sequential swaps of an index vector
    idx = np.arange(n)
    for i, p in enumerate(piv):
        if p > 0:
            idx[[i, i + p]] = idx[[i + p, i]]
    np.testing.assert_array_equal(X, X0[idx])
    np.testing.assert_array_equal(O.laswp(X, piv, forward=False), X0)
    Y = O.laswp(X0.T.copy(), piv, side=O.RIGHT)
    np.testing.assert_array_equal(Y, X0[idx].T)


def test_getrf_nopiv_vs_reconstruction():
    n = 300
    L = np.tril(O.uniform_matrix(n, n, 1), -1) * 0.1 + np.eye(n)
    U = np.triu(O.uniform_matrix(n, n, 2)) + 4 * np.eye(n)
    A = L @ U
    assert O.getrf_nopiv(A) == 0
    Lf, Uf = O.lu_factors(A)
    assert rel_fro(Lf, L) < 1e-12 and rel_fro(Uf, U) < 1e-12


# -------------------------------------------------------------------------- Cholesky ---
def test_potrf_integer_known_answers():
    c = golden("chol_int.npz")
    for t in range(int(c["count"])):
        A = c[f"A{t}"].copy()
        assert O.potrf(A) == 0
        np.testing.assert_array_equal(np.tril(A), c[f"L{t}"])
        np.testing.assert_array_equal(np.triu(A, 1), np.triu(c[f"A{t}"], 1))  # untouched
        U = c[f"A{t}"].copy()
        assert O.potrf(U, O.UPPER) == 0
        np.testing.assert_array_equal(np.triu(U), c[f"L{t}"].T)


@pytest.mark.parametrize("n", [7, 128, 129, 300])
def test_potrf_vs_lapack_and_solve(n):
    G = O.uniform_matrix(n, n, 50 + n)
    S = G @ G.T + n * np.eye(n)
    L = S.copy()
    assert O.potrf(L) == 0
    ref = np.linalg.cholesky(S)
    assert rel_fro(np.tril(L), ref) < 1e-14 * n
    B = O.uniform_matrix(n, 5, 3)
    X = O.potrs(L, B.copy())
    assert rel_fro(S @ X, B) < 1e-12
    U = S.copy()
    assert O.potrf(U, O.UPPER) == 0
    X = O.potrs(U, B.copy(), O.UPPER)
    assert rel_fro(S @ X, B) < 1e-12


def test_potrf_not_spd_info():
    S = np.eye(200)
    S[150, 150] = -1.0
    assert O.potrf(S.copy()) == 151
    assert O.potrf(S.copy(), O.UPPER) == 151


# ------------------------------------------------------------------------ TRSM / SYRK ---
@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("uplo", [0xC0, 0x60])
@pytest.mark.parametrize("trans", [0, 8])
@pytest.mark.parametrize("diag", [0, 0x100])
def test_trsm_all_variants(side, uplo, trans, diag):
    m, n = 150, 37
    k = m if side == 0 else n
    A = O.uniform_matrix(k, k, 5) * 0.1 + 3 * np.eye(k)
    B = O.uniform_matrix(m, n, 6)
    X = O.trsm(side, uplo, trans, diag, 2.0, A, B.copy())
    T = np.tril(A) if uplo == 0xC0 else np.triu(A)
    if diag == 0x100:
        np.fill_diagonal(T, 1.0)
    opT = T.T if trans == 8 else T
    lhs = opT @ X if side == 0 else X @ opT
    assert rel_fro(lhs, 2.0 * B) < 1e-13


@pytest.mark.parametrize("uplo", [0xC0, 0x60])
@pytest.mark.parametrize("trans", [0, 8])
def test_syrk(uplo, trans):
    n, k = 140, 33
    A = O.uniform_matrix(*(k, n) if trans else (n, k), 9)
    C0 = O.uniform_matrix(n, n, 10)
    C = C0.copy()
    O.syrk(uplo, trans, -1.0, A, 0.5, C)
    opA = A.T if trans else A
    full = 0.5 * C0 - opA @ opA.T
    mask = np.tril(np.ones((n, n), bool)) if uplo == 0xC0 else np.triu(np.ones((n, n), bool))
    assert rel_fro(C[mask], full[mask]) < 1e-14 * k
    np.testing.assert_array_equal(C[~mask], C0[~mask])


def test_uniform_stream_closed_form():
    # x_i = (splitmix64(seed + (i+1)*golden) >> 11) * 2^-52 - 1, checked in pure Python
    def splitmix(x):
        M = 2**64 - 1
        x ^= x >> 30
        x = (x * 0xBF58476D1CE4E5B9) & M
        x ^= x >> 27
        x = (x * 0x94D049BB133111EB) & M
        x ^= x >> 31
        return x

    seed = 0xDD1A0001
    x = O.uniform(5, seed)
    for i in range(5):
        z = splitmix((seed + (i + 1) * 0x9E3779B97F4A7C15) & (2**64 - 1))
        assert x[i] == (z >> 11) * 2.0**-52 - 1.0
    y = O.uniform(100000, 7)
    assert -1.0 <= y.min() and y.max() < 1.0 and abs(y.mean()) < 0.01
