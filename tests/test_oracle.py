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


# ------------------------------------------------------------------ tournament pivoting ---
def designed_panel(m, seed, off=0.0):
    """m x 128 panel whose 128 pivot rows are known: rows e_j (1 in column j, `off`-sized
    entries after column j: one unit row of a unit lower trapezoid) at random places, every
    other row zero, so partial pivoting and the tournament must pick e_0, e_1, ... in order."""
    rng = np.random.default_rng(seed)
    P = np.zeros((m, 128), dtype=np.float32)
    where = rng.permutation(m)[:128]
    for j, r in enumerate(where):
        P[r, j] = 1.0
        P[r, j + 1:] = off * rng.uniform(-1, 1, 127 - j)
    return P, where


def check_tslu_structure(P, zero, sel, top, pairs):
    m = P.shape[0]
    assert len(set(sel.tolist())) == 128 and sel.min() >= 0 and sel.max() < m
    # L11 U11 reproduces the selected rows (FP32 elimination)
    L = np.tril(top.astype(np.float64), -1) + np.eye(128)
    U = np.triu(top.astype(np.float64))
    A = P[sel].astype(np.float64)
    assert np.linalg.norm(L @ U - A) <= 1e-5 * np.linalg.norm(A)
    # the pairs are one permutation that brings sel to the top
    src = np.arange(m)
    new = src.copy()
    new[pairs[:, 0]] = src[pairs[:, 1]]
    assert sorted(new.tolist()) == list(range(m))
    np.testing.assert_array_equal(new[:128], sel)


@pytest.mark.parametrize("m", [128, 200, 256, 700, 1300, 2100])
def test_tslu_designed_pivot_order(oracle, m):
    # every leaf keeps the unit rows it holds and every merge passes them on (they hold 1
    # where all other candidates hold 0), so the root eliminates them in column order;
    # 1300 / 2100 rows: trees with odd levels (6, 3, 2, 1 and 9, 5, 3, 2, 1 nodes)
    P, where = designed_panel(m, m)
    zero, sel, top, pairs = oracle.tslu(P)
    assert zero == 0
    np.testing.assert_array_equal(sel, where)
    check_tslu_structure(P, zero, sel, top, pairs)
    # partial pivoting (the reference's panel rule) picks the same rows
    _, piv = oracle.getf2(P.astype(np.float64).copy())
    order = np.arange(m)
    for i, p in enumerate(piv):
        order[[i, i + p]] = order[[i + p, i]]
    np.testing.assert_array_equal(order[:128], where)


@pytest.mark.parametrize("m", [150, 256])
def test_tslu_single_stack_is_partial_pivoting(oracle, m):
    # one stack: the tournament is GEPP of the panel.  P = L0 U0 with L0 = unit rows at known
    # places and |l| < 0.45 elsewhere: eliminating with a unit row changes no later column
    # of L0, so each unit row wins its column by more than a factor of two, and right
    # multiplication by the upper triangular U0 keeps that choice.  The reference panel
    # (partial pivoting) picks the same rows and its factored top equals the tournament's
    # to FP32 accuracy.
    L0, where = designed_panel(m, 7 + m)
    rng = np.random.default_rng(m)
    rest = np.setdiff1d(np.arange(m), where)
    L0 = L0.astype(np.float64)
    L0[rest] = 0.45 * rng.uniform(-1, 1, (rest.size, 128))
    U0 = np.triu(rng.uniform(-1, 1, (128, 128)), 1) + np.diag(rng.uniform(1, 2, 128))
    P = (L0 @ U0).astype(np.float32)
    zero, sel, top, pairs = oracle.tslu(P)
    assert zero == 0
    np.testing.assert_array_equal(sel, where)
    ref = P.astype(np.float64)
    _, piv = oracle.getf2(ref)
    order = np.arange(m)
    for i, p in enumerate(piv):
        order[[i, i + p]] = order[[i + p, i]]
    np.testing.assert_array_equal(order[:128], where)
    np.testing.assert_allclose(top, ref[:128], rtol=0, atol=1e-5 * np.abs(ref[:128]).max())
    check_tslu_structure(P, zero, sel, top, pairs)


@pytest.mark.parametrize("m", [300, 1024, 4000])
def test_tslu_first_pivot_is_column_max_and_structure(oracle, m):
    # the global arg-max of column 0 wins its stack and every merge: the first pivot equals
    # partial pivoting's; the rest is checked structurally
    P = np.random.default_rng(m).uniform(-1, 1, (m, 128)).astype(np.float32)
    zero, sel, top, pairs = oracle.tslu(P)
    assert zero == 0
    assert sel[0] == int(np.argmax(np.abs(P[:, 0])))
    check_tslu_structure(P, zero, sel, top, pairs)
    # multipliers of the selected rows stay bounded (pivot growth close to GEPP's)
    assert np.abs(np.tril(top, -1)).max() <= 1.0 + 1e-6


def test_tslu_dominant_top_block_keeps_order(oracle):
    # a strongly diagonally dominant top block over tiny rows: no interchange at all
    rng = np.random.default_rng(5)
    m = 1500
    P = (1e-3 * rng.uniform(-1, 1, (m, 128))).astype(np.float32)
    P[:128] += (10 * np.eye(128) + 0.05 * rng.uniform(-1, 1, (128, 128))).astype(np.float32)
    zero, sel, top, pairs = oracle.tslu(P)
    np.testing.assert_array_equal(sel, np.arange(128))
    assert pairs.shape[0] == 0
    check_tslu_structure(P, zero, sel, top, pairs)


def test_tslu_zero_column_reports_info(oracle):
    P = np.random.default_rng(2).uniform(-1, 1, (600, 128)).astype(np.float32)
    P[:, 40] = 0.0
    zero, sel, top, pairs = oracle.tslu(P)
    assert zero == 41


# ---------------------------------------------------------------- level-3 BLAS semantics ---
# The oracle's GEMMT / SYR2K / SYMM / TRMM / TRMM3 (BLIS semantics over the restated managed
# GEMM) against scipy's BLAS, an independent implementation of the same definitions.
import scipy.linalg.blas as sblas  # noqa: E402

LO, UP = 0xC0, 0x60


@pytest.mark.parametrize("uplo", [LO, UP])
@pytest.mark.parametrize("ta,tb", [(0, 0), (8, 0), (0, 8), (8, 8)])
def test_oracle_gemmt_triangle(uplo, ta, tb):
    m, k = 37, 150
    A = O.uniform_matrix(*(k, m) if ta else (m, k), 1)
    B = O.uniform_matrix(*(m, k) if tb else (k, m), 2)
    C0 = O.uniform_matrix(m, m, 3)
    C = O.gemmt(uplo, ta, tb, 0.7, A, B, 1.3, C0.copy())
    full = 0.7 * ((A.T if ta else A) @ (B.T if tb else B)) + 1.3 * C0
    mask = np.tril(np.ones((m, m), bool)) if uplo == LO else np.triu(np.ones((m, m), bool))
    np.testing.assert_allclose(C[mask], full[mask], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(C[~mask], C0[~mask])


@pytest.mark.parametrize("uplo", [LO, UP])
@pytest.mark.parametrize("trans", [0, 8])
def test_oracle_syr2k_vs_blas(uplo, trans):
    m, k = 41, 90
    A = O.uniform_matrix(*(k, m) if trans else (m, k), 4)
    B = O.uniform_matrix(*(k, m) if trans else (m, k), 5)
    C0 = O.uniform_matrix(m, m, 6)
    C = O.syr2k(uplo, trans, trans, 1.1, A, B, 0.4, C0.copy())
    ref = sblas.dsyr2k(1.1, A, B, beta=0.4, c=C0.copy(), trans=1 if trans else 0,
                       lower=1 if uplo == LO else 0)
    mask = np.tril(np.ones((m, m), bool)) if uplo == LO else np.triu(np.ones((m, m), bool))
    np.testing.assert_allclose(C[mask], ref[mask], rtol=0, atol=1e-12)


@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("uplo", [LO, UP])
def test_oracle_symm_vs_blas(side, uplo):
    m, n = 45, 30
    k = m if side == 0 else n
    A = O.uniform_matrix(k, k, 7)  # only the uplo triangle may be read
    B = O.uniform_matrix(m, n, 8)
    C0 = O.uniform_matrix(m, n, 9)
    C = O.symm(side, uplo, 0, 1 / np.pi, A, B, np.e, C0.copy())
    ref = sblas.dsymm(1 / np.pi, A, B, beta=np.e, c=C0.copy(), side=side,
                      lower=1 if uplo == LO else 0)
    np.testing.assert_allclose(C, ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("uplo", [LO, UP])
@pytest.mark.parametrize("trans", [0, 8])
@pytest.mark.parametrize("diag", [0x000, 0x100])
def test_oracle_trmm_vs_blas(side, uplo, trans, diag):
    m, n = 33, 52
    k = m if side == 0 else n
    A = O.uniform_matrix(k, k, 10)
    B0 = O.uniform_matrix(m, n, 11)
    B = O.trmm(side, uplo, trans, diag, 1.2, A, B0.copy())
    ref = sblas.dtrmm(1.2, A, B0.copy(), side=side, lower=1 if uplo == LO else 0,
                      trans_a=1 if trans else 0, diag=1 if diag == 0x100 else 0)
    np.testing.assert_allclose(B, ref, rtol=0, atol=1e-12)
    # TRMM3 with beta = 0 and op(B) = B is the same product
    C = O.trmm3(side, uplo, trans, diag, 0, 1.2, A, B0, 0.0, np.full((m, n), np.nan))
    np.testing.assert_allclose(C, ref, rtol=0, atol=1e-12)


# ---------------------------------------------------------------- LDL^T restatement ----
# Pinned by: a hand-computed 3 x 3 factor; the identity L sqrt(D) = chol(A) for SPD A
# (numpy's LAPACK Cholesky); reconstruction L D L^T = A for a symmetric indefinite matrix;
# the blocked path (n > 128) against the pure-loop unblocked path on the same matrix; the
# zero-pivot index of a designed matrix; and solves against numpy.linalg.solve.

def _sym_indefinite(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.uniform(-1, 1, (n, n))
    S = (M + M.T) / 2
    # diagonally dominant with mixed signs: every leading minor is far from zero
    S[np.diag_indices(n)] = (n / 2 + rng.uniform(0, 1, n)) * np.where(np.arange(n) % 3 == 1, -1, 1)
    return S


def test_ldlt_oracle_hand_computed():
    A = np.array([[4.0, 2.0, -2.0], [2.0, 5.0, 1.0], [-2.0, 1.0, 3.0]])
    # d1 = 4, l21 = 0.5, l31 = -0.5; d2 = 5 - 4 (0.25) = 4, l32 = (1 - 4 (0.5)(-0.5)) / 4 = 0.5;
    # d3 = 3 - 4 (0.25) - 4 (0.25) = 1
    F = A.copy()
    assert O.ldlt(F) == 0
    L, d = O.ldlt_factors(F)
    np.testing.assert_array_equal(d, [4.0, 4.0, 1.0])
    np.testing.assert_array_equal(L, [[1, 0, 0], [0.5, 1, 0], [-0.5, 0.5, 1]])
    np.testing.assert_array_equal(np.triu(F, 1), np.triu(A, 1))  # other triangle untouched
    Fu = A.copy()
    assert O.ldlt(Fu, O.UPPER) == 0
    U, du = O.ldlt_factors(Fu, O.UPPER)
    np.testing.assert_array_equal(du, d)
    np.testing.assert_array_equal(U, L.T)


@pytest.mark.parametrize("n", [7, 128, 300])
def test_ldlt_oracle_matches_cholesky_on_spd(n):
    rng = np.random.default_rng(n)
    M = rng.uniform(-1, 1, (n, n))
    A = M @ M.T + n * np.eye(n)
    for uplo in (O.LOWER, O.UPPER):
        F = A.copy()
        assert O.ldlt(F, uplo) == 0
        T, d = O.ldlt_factors(F, uplo)
        Lc = np.linalg.cholesky(A)
        G = T * np.sqrt(d)[None, :] if uplo == O.LOWER else (np.sqrt(d)[:, None] * T).T
        assert np.max(np.abs(G - Lc)) < 1e-12 * np.max(np.abs(Lc))


@pytest.mark.parametrize("n", [5, 130, 300])
def test_ldlt_oracle_reconstructs_indefinite_and_blocked_equals_unblocked(n):
    A = _sym_indefinite(n, 1000 + n)
    F = A.copy()
    assert O.ldlt(F) == 0
    L, d = O.ldlt_factors(F)
    assert np.linalg.norm(L @ np.diag(d) @ L.T - A) < 1e-13 * n * np.linalg.norm(A)
    assert (d < 0).any() and (d > 0).any()
    Fu = A.copy()
    assert O.ldlt(Fu, block=10 ** 6) == 0  # one unblocked pass over the whole matrix
    assert np.max(np.abs(np.tril(Fu) - np.tril(F))) < 1e-12


def test_ldlt_oracle_zero_pivot_index():
    A = _sym_indefinite(10, 5)
    # row and column 3 decoupled from the first three pivots and a zero diagonal: the first
    # three steps leave a_33 = 0, so d_3 = 0 exactly
    A[3, :4] = 0.0
    A[:4, 3] = 0.0
    assert O.ldlt(A.copy()) == 4
    # the last pivot of a block is not tested (LDLT.cs:145 stops at Rows - 1): no status
    B = _sym_indefinite(6, 9)
    B[5, :] = 0.0
    B[:, 5] = 0.0
    assert O.ldlt(B.copy()) == 0


def test_ldlt_oracle_solve():
    A = _sym_indefinite(150, 7)
    B = np.random.default_rng(8).uniform(-1, 1, (150, 3))
    for uplo in (O.LOWER, O.UPPER):
        F = A.copy()
        assert O.ldlt(F, uplo) == 0
        X = O.ldlt_solve(F, B.copy(), uplo)
        np.testing.assert_allclose(X, np.linalg.solve(A, B), rtol=1e-10, atol=1e-12)


# ------------------------------------------------------------------ QR restatement ----
# Pinned by LAPACK's dgeqrf (scipy mode='raw'), which uses the same Householder convention
# (beta = -sign(alpha) |x|, v = x / (alpha - beta), v_1 = 1): R and the vectors agree, and the
# reference's tau is LAPACK's 1/tau.  Where a column has no entries below the diagonal
# (the last one when m <= n) LAPACK leaves H = I (tau = 0) while the reference reflects
# (tau = 1/2, the row of R negated).  T is pinned by its definition (UT transform:
# T = strictly-upper(U^T U) + diag(tau)); Q by orthogonality and A = Q R; the solve by
# numpy's least squares.

@pytest.mark.parametrize("m,n", [(5, 3), (300, 300), (700, 520), (100, 300), (3, 5), (600, 257)])
def test_qr_oracle_matches_lapack(m, n):
    A0 = np.random.default_rng(m * 7 + n).uniform(-1, 1, (m, n))
    A = A0.copy()
    T = O.qr(A)
    assert T.shape == (min(256, max(m, n)), n)
    (h, tau), _ = sla.qr(A0, mode="raw")
    k = min(m, n)
    full = m > n  # every column has entries below its diagonal
    rows = m if full else m - 1
    assert np.max(np.abs(np.triu(A)[:rows] - np.triu(h)[:rows])) < 1e-13
    if not full:
        np.testing.assert_allclose(A[m - 1, m - 1:], -h[m - 1, m - 1:], rtol=0, atol=1e-13)
    assert np.max(np.abs(np.tril(A, -1)[:, :k] - np.tril(h, -1)[:, :k])) < 1e-13
    taus = np.array([T[j % 256, j] for j in range(k)])
    kk = k if full else k - 1
    np.testing.assert_allclose(taus[:kk], 1.0 / tau[:kk], rtol=1e-13)
    if not full:
        assert taus[k - 1] == 0.5
    # T: strictly upper part of U^T U per block, tau on the diagonal
    for k0 in range(0, k, 256 if n > 256 else max(k, 1)):
        b = min(256 if n > 256 else k, k - k0)
        U = O._unit_lower(A[k0:, k0:k0 + b])
        G = U.T @ U
        Tb = T[:b, k0:k0 + b]
        assert np.max(np.abs(np.triu(Tb, 1) - np.triu(G, 1))) < 1e-12 * max(1.0, np.max(np.abs(G)))
    Q = O.qr_apply(A, T, np.eye(m), False)
    assert np.max(np.abs(Q.T @ Q - np.eye(m))) < 1e-13
    assert np.max(np.abs(Q @ np.triu(A) - A0)) < 1e-13
    Qt = O.qr_apply(A, T, np.eye(m), True)
    assert np.max(np.abs(Qt - Q.T)) < 1e-13


@pytest.mark.parametrize("m,n,nrhs", [(40, 40, 1), (500, 300, 3), (800, 600, 2)])
def test_qr_oracle_least_squares(m, n, nrhs):
    rng = np.random.default_rng(m + n)
    A0 = rng.uniform(-1, 1, (m, n))
    B = rng.uniform(-1, 1, (m, nrhs))
    A = A0.copy()
    T = O.qr(A)
    X = O.qr_solve(A, T, B)
    Xl = np.linalg.lstsq(A0, B, rcond=None)[0]
    np.testing.assert_allclose(X, Xl, rtol=1e-9, atol=1e-11)


# ------------------------------------------------------- tensor contraction restatement ----
# Pinned by numpy.einsum wherever einsum can state the same contraction (products, batches,
# transposed outputs, diagonals of inputs, indices summed in one operand), and by hand-built
# answers for what einsum cannot state (an index of the result alone is a broadcast, a
# repeated index of the result its diagonal).

@pytest.mark.parametrize("expr,sa,sb", [
    ("ij,jk->ik", (4, 5), (5, 3)),
    ("ij,jk->ki", (4, 5), (5, 3)),
    ("bij,bjk->bik", (3, 4, 5), (3, 5, 2)),
    ("abcd,cdef->abef", (2, 3, 4, 2), (4, 2, 3, 2)),
    ("iij,jk->ik", (4, 4, 3), (3, 2)),
    ("ijx,jk->ik", (3, 4, 2), (4, 5)),
    ("ij,kl->ijkl", (2, 3), (2, 2)),
    ("i,i->", (7,), (7,)),
])
def test_contract_oracle_matches_einsum(expr, sa, sb):
    rng = np.random.default_rng(len(expr))
    A = rng.uniform(-1, 1, sa)
    B = rng.uniform(-1, 1, sb)
    ref = np.einsum(expr, A, B)
    C0 = rng.uniform(-1, 1, ref.shape)
    C = C0.copy()
    O.contract(expr, 0.75, A, B, 0.5, C)
    np.testing.assert_allclose(C, 0.5 * C0 + 0.75 * ref, rtol=1e-13, atol=1e-14)


def test_contract_oracle_strided_views_broadcast_and_result_diagonal():
    rng = np.random.default_rng(5)
    big = rng.uniform(-1, 1, (6, 10))
    A = big[:, ::2]          # 6 x 5 view, column stride 2
    B = rng.uniform(-1, 1, (3, 5)).T  # 5 x 3 transposed view
    C = np.zeros((6, 3, 2))
    O.contract("ij,jk->ikz", 1.0, A, B, 0.0, C)  # z only in C: broadcast
    for z in range(2):
        np.testing.assert_allclose(C[:, :, z], A @ B, rtol=1e-13)
    D = np.zeros((4, 4))
    X = rng.uniform(-1, 1, (4, 3))
    Y = rng.uniform(-1, 1, (3, 4))
    O.contract("ij,ji->ii", 1.0, X, Y, 0.0, D)  # a repeated result index: its diagonal
    np.testing.assert_allclose(np.diag(D), np.einsum("ij,ji->i", X, Y), rtol=1e-13)
    assert np.count_nonzero(D - np.diag(np.diag(D))) == 0
    with pytest.raises(ValueError):
        O.contract("ij,jk", 1.0, X, Y, 0.0, D)
    with pytest.raises(ValueError):
        O.contract("ij,jk->ik", 1.0, X, rng.uniform(-1, 1, (4, 4)), 0.0, D)
