"""GPU parity of the blocked LDL^T (no pivoting) and its solve against the CPU oracle's
restatement of LDLT.cs: factors within 1e-12 relative (max-abs) error, the other triangle
untouched bit for bit, zero pivots reported with their index."""
import numpy as np
import pytest

from test_oracle import _sym_indefinite

pytestmark = pytest.mark.gpu


def _spd(n, seed):
    M = np.random.default_rng(seed).uniform(-1, 1, (n, n))
    return M @ M.T + n * np.eye(n)


def _tri(X, uplo, lower_code):
    return np.tril(X) if uplo == lower_code else np.triu(X)


@pytest.mark.parametrize("n", [1, 2, 64, 127, 128, 129, 300, 1000])
@pytest.mark.parametrize("kind", ["spd", "indefinite"])
def test_ldlt_matches_oracle(ddla, oracle, n, kind):
    S = _spd(n, n) if kind == "spd" else _sym_indefinite(n, 2000 + n)
    junk = np.full((n, n), 7.25)  # the triangle not named must be neither read nor written
    for uplo, ou in ((ddla.LOWER, oracle.LOWER), (ddla.UPPER, oracle.UPPER)):
        if uplo == ddla.LOWER:
            src = np.tril(S) + np.triu(junk, 1)
        else:
            src = np.triu(S) + np.tril(junk, -1)
        Fo = src.copy()
        assert oracle.ldlt(Fo, ou) == 0
        d = ddla.DeviceMatrix.from_host(src)
        ddla.LDLTDecompose(uplo, d)
        F = d.to_host()
        mine, ref = _tri(F, uplo, ddla.LOWER), _tri(Fo, uplo, ddla.LOWER)
        assert np.max(np.abs(mine - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
        np.testing.assert_array_equal(F - mine, src - _tri(src, uplo, ddla.LOWER))


@pytest.mark.parametrize("n,nrhs", [(129, 1), (500, 3), (1000, 16)])
def test_ldlt_solve_matches_oracle(ddla, oracle, n, nrhs):
    A = _sym_indefinite(n, 3000 + n)
    B = np.random.default_rng(n).uniform(-1, 1, (n, nrhs))
    for uplo in (ddla.LOWER, ddla.UPPER):
        ou = oracle.LOWER if uplo == ddla.LOWER else oracle.UPPER
        Fo = A.copy()
        assert oracle.ldlt(Fo, ou) == 0
        Xo = oracle.ldlt_solve(Fo, B.copy(), ou)
        d = ddla.DeviceMatrix.from_host(A)
        ddla.LDLTDecompose(uplo, d)
        db = ddla.DeviceMatrix.from_host(B)
        ddla.LDLTSolve(uplo, d, db)
        X = db.to_host()
        assert np.max(np.abs(X - Xo)) <= 1e-11 * np.max(np.abs(Xo))
        r = B - A @ X
        assert np.linalg.norm(r, np.inf) / (np.linalg.norm(A, np.inf) * np.linalg.norm(X, np.inf)) < 1e-14


def test_ldlt_host_strided_and_class(ddla, oracle):
    n = 260
    A = _sym_indefinite(n, 41)
    # column-major view with column stride 1 and row stride n (Fortran order)
    Af = np.asfortranarray(A.copy())
    ddla.LDLTDecompose(ddla.LOWER, Af)
    Fo = A.copy()
    oracle.ldlt(Fo)
    assert np.max(np.abs(np.tril(Af) - np.tril(Fo))) <= 1e-12 * np.max(np.abs(Fo))
    # a view with column stride 2
    big = np.zeros((n, 2 * n))
    big[:, ::2] = A
    V = big[:, ::2]
    ddla.LDLTDecompose(ddla.UPPER, V)
    Fu = A.copy()
    oracle.ldlt(Fu, oracle.UPPER)
    assert np.max(np.abs(np.triu(V) - np.triu(Fu))) <= 1e-12 * np.max(np.abs(Fu))
    np.testing.assert_array_equal(big[:, 1::2], 0.0)
    # the class: factor once, solve, deconstruct
    f = ddla.LDLT(A)
    b = np.random.default_rng(3).uniform(-1, 1, n)
    x = f.Solve(b)
    np.testing.assert_allclose(A @ x, b, rtol=0, atol=1e-11)
    L, D = f.Deconstruct()
    np.testing.assert_allclose(L @ D @ L.T, A, rtol=0, atol=1e-11 * np.max(np.abs(A)))
    with pytest.raises(RuntimeError):
        f.Solve(b)


def test_ldlt_zero_pivot(ddla, oracle):
    # a zero pivot inside a block: the same index as the reference
    A = _sym_indefinite(300, 5)
    A[3, :4] = 0.0  # row and column 3 decoupled from the first pivots, d_3 = 0 exactly
    A[:4, 3] = 0.0
    assert oracle.ldlt(A.copy()) == 4
    with pytest.raises(ddla.LinalgException, match="at 3"):
        ddla.LDLTDecompose(ddla.LOWER, A.copy())
    # the last pivot of the first 128-block: the reference does not test it, this library does
    B = _sym_indefinite(300, 6)
    B[127, :128] = 0.0
    B[:128, 127] = 0.0
    d = ddla.DeviceMatrix.from_host(B)
    st = ddla.native.lib().ddla_dldlt_dev(ddla.LOWER, 300, d.ptr, d.ld)
    assert st == 128
    # argument errors
    assert ddla.native.lib().ddla_dldlt_dev(0x11, 300, d.ptr, d.ld) == ddla.native.ERR_INVALID_ARG
