"""GPU parity of the FP64 blocked LU (partial pivoting) and Cholesky against the CPU oracle:
the pivot sequence must match exactly; factors within 1e-13*n relative Frobenius error."""
import numpy as np
import pytest

from conftest import golden, rel_fro

pytestmark = pytest.mark.gpu

SEED_C2 = 0xDD1A0002


def backward_error(A, x, b):
    r = b - A @ x
    return np.linalg.norm(r, np.inf) / (np.linalg.norm(A, np.inf) * np.linalg.norm(x, np.inf))


def test_lu_exact_gepp_fixtures(ddla):
    e = golden("gepp_exact.npz")
    for t in range(int(e["count"])):
        A = e[f"A{t}"].copy()
        piv = ddla.PLUDecompose(A)
        np.testing.assert_array_equal(piv, e[f"piv{t}"])
        np.testing.assert_allclose(A, e[f"LU{t}"], rtol=0, atol=1e-12)


def test_lu_designed_known_answers(ddla):
    d = golden("lu_designed.npz")
    for t in range(int(d["count"])):
        A = d[f"A{t}"].copy()
        piv = ddla.PLUDecompose(A)
        np.testing.assert_array_equal(piv, d[f"piv{t}"])
        np.testing.assert_array_equal(np.tril(A, -1) + np.eye(len(A)), d[f"L{t}"])
        np.testing.assert_array_equal(np.triu(A), d[f"U{t}"])


@pytest.mark.parametrize("n", [1, 5, 64, 127, 128, 129, 200, 257, 513, 1000, 2048])
def test_lu_pivots_and_factors_match_oracle(ddla, oracle, n):
    A0 = oracle.uniform_matrix(n, n, SEED_C2 + n)
    Ao = A0.copy()
    info, piv_o = oracle.getrf(Ao)
    assert info == 0
    A = A0.copy()
    piv = ddla.PLUDecompose(A)
    np.testing.assert_array_equal(piv, piv_o)
    assert rel_fro(A, Ao) <= 1e-13 * n
    # reconstruction P A = L U (LU.Deconstruct, TestPLU.cs:79-116)
    L = np.tril(A, -1) + np.eye(n)
    U = np.triu(A)
    P = oracle.perm_matrix(piv)
    assert rel_fro(P @ A0, L @ U) <= 1e-14 * n


def test_lu_device_resident_and_solve(ddla, oracle):
    n, nrhs = 777, 3
    A0 = oracle.uniform_matrix(n, n, 42)
    B0 = oracle.uniform_matrix(n, nrhs, 43)
    dA = ddla.DeviceMatrix.from_host(A0)
    piv = ddla.PLUDecompose(dA)
    Ao = A0.copy()
    _, piv_o = oracle.getrf(Ao)
    np.testing.assert_array_equal(piv, piv_o)
    for trans in (False, True):
        dB = ddla.DeviceMatrix.from_host(B0)
        ddla.PLUSolve(dA, piv, dB, trans)
        X = dB.to_host()
        Xo = oracle.getrs(Ao, piv_o, B0.copy(), trans=trans)
        assert rel_fro(X, Xo) <= 1e-11
        M = A0.T if trans else A0
        for j in range(nrhs):
            assert backward_error(M, X[:, j], B0[:, j]) < 1e-15 * 10


def test_lu_class_mirror(ddla):
    n = 300
    rng = np.random.default_rng(5)
    A = rng.uniform(-1, 1, (n, n))
    b = rng.uniform(-1, 1, n)
    lu = ddla.LU(A)
    x = lu.Solve(b)
    assert backward_error(A, x, b) < 1e-14
    P, L, U = lu.Deconstruct()
    assert rel_fro(P @ A, L @ U) < 1e-13


@pytest.mark.parametrize("n", [64, 300, 700])
def test_lu_class_matches_oracle(ddla, oracle, n):
    # the LU class (LU.cs:13-128) against the oracle: pivots identical, factors, both solves
    # and the inverse (the reference's Inverse returns the identity; the mirror's is A^-1)
    A0 = oracle.uniform_matrix(n, n, 0xDD1A0100 + n)
    Ao = A0.copy()
    info, piv_o = oracle.getrf(Ao)
    assert info == 0
    lu = ddla.LU(A0)
    B = oracle.uniform_matrix(n, 3, 7)
    for trans in (False, True):
        x = lu.Solve(B, trans=trans)
        xo = oracle.getrs(Ao, piv_o, B.copy(), trans=trans)
        assert rel_fro(x, xo) < 1e-12
    np.testing.assert_array_equal(lu.pivs, piv_o)
    assert rel_fro(lu.matrix, Ao) < 1e-13 * n
    inv = lu.Inverse()
    assert rel_fro(inv, np.linalg.inv(A0)) < 1e-11
    assert np.max(np.abs(A0 @ inv - np.eye(n))) < 1e-11
    P, L, U = lu.Deconstruct()
    Lo, Uo = oracle.lu_factors(Ao)
    np.testing.assert_array_equal(P, oracle.perm_matrix(piv_o))
    assert rel_fro(L, Lo) < 1e-13 * n and rel_fro(U, Uo) < 1e-13 * n


def test_lu_zero_pivot_raises(ddla):
    A = np.random.default_rng(1).uniform(-1, 1, (200, 200))
    A[:, 150] = 0.0
    with pytest.raises(ZeroDivisionError):
        ddla.PLUDecompose(A)
    # info value through the C ABI: first zero pivot, 1-based
    d = ddla.DeviceMatrix.from_host(np.random.default_rng(1).uniform(-1, 1, (200, 200)) * 0 +
                                    np.diag(np.r_[np.ones(150), 0.0, np.ones(49)]))
    dpiv = ddla._DevInts(200)
    assert ddla.native.lib().ddla_dgetrf_dev(200, d.ptr, d.ld, dpiv.ptr) == 151


def test_laswp_device(ddla, oracle):
    n = 300
    rng = np.random.default_rng(2)
    piv = np.array([rng.integers(0, n - i) for i in range(n)], dtype=np.int32)
    X0 = oracle.uniform_matrix(n, 20, 4)
    for forward in (True, False):
        d = ddla.DeviceMatrix.from_host(X0)
        ddla.ApplyPiv(d, piv, forward=forward)
        np.testing.assert_array_equal(d.to_host(), oracle.laswp(X0.copy(), piv, forward=forward))
    Y0 = oracle.uniform_matrix(20, n, 5)
    d = ddla.DeviceMatrix.from_host(Y0)
    ddla.ApplyPiv(d, piv, side=ddla.RIGHT)
    np.testing.assert_array_equal(d.to_host(), oracle.laswp(Y0.copy(), piv, side=oracle.RIGHT))


def test_laswp_device_pivots_across_blocks(ddla, oracle):
    # the pivots are turned into row moves per block of 128 and then composed: runs of rows
    # that all exchange with one far row, partners several blocks ahead, a short last block
    n = 1000
    rng = np.random.default_rng(7)
    piv = np.zeros(n, dtype=np.int32)
    for i in range(n):
        r = rng.random()
        if r < 0.3:
            piv[i] = max(0, 900 - i)          # many rows with the same partner
        elif r < 0.6:
            piv[i] = rng.integers(0, n - i)   # anywhere below
        elif r < 0.8:
            piv[i] = min(n - 1 - i, 1)        # the next row
    X0 = oracle.uniform_matrix(n, 8, 6)
    for forward in (True, False):
        d = ddla.DeviceMatrix.from_host(X0)
        ddla.ApplyPiv(d, piv, forward=forward)
        np.testing.assert_array_equal(d.to_host(), oracle.laswp(X0.copy(), piv, forward=forward))


@pytest.mark.slow
def test_lu_8192_pivot_sequence_and_backward_error(ddla, oracle):
    # config: n=8192, A uniform[-1,1], b = A x_true; pivot sequence must match exactly
    n = 8192
    A0 = oracle.uniform_matrix(n, n, SEED_C2)
    xt = oracle.uniform(n, SEED_C2 + 1).reshape(-1, 1)
    b = oracle.matmul(A0, xt)
    Ao = A0.copy()
    info, piv_o = oracle.getrf(Ao)
    assert info == 0
    dA = ddla.DeviceMatrix.from_host(A0)
    piv = ddla.PLUDecompose(dA)
    mism = np.nonzero(piv != piv_o)[0]
    assert mism.size == 0, f"first divergent pivot at step {mism[0] if mism.size else -1}"
    db = ddla.DeviceMatrix.from_host(b)
    ddla.PLUSolve(dA, piv, db)
    x = db.to_host()
    assert backward_error(A0, x[:, 0], b[:, 0]) < 1e-15 * 20
    assert rel_fro(dA.to_host(), Ao) <= 1e-13 * n


# --------------------------------------------------------------- LU without pivoting ---
@pytest.mark.parametrize("n", [1, 7, 128, 129, 300, 1000])
def test_lu_nopiv_matches_oracle(ddla, oracle, n):
    # diagonally dominant, so elimination without pivoting is stable (TestLU.cs:11-63 builds
    # A = L U from random factors the same way)
    L = np.tril(oracle.uniform_matrix(n, n, 3 + n), -1) * 0.1 + np.eye(n)
    U = np.triu(oracle.uniform_matrix(n, n, 5 + n)) + 4 * np.eye(n)
    A0 = L @ U
    Ao = A0.copy()
    assert oracle.getrf_nopiv(Ao) == 0
    A = A0.copy()
    ddla.LUDecompose(A)
    assert rel_fro(A, Ao) <= 1e-13 * n
    # A0 = L U was rounded, so its factors differ from the generators by about cond * u
    Lf, Uf = oracle.lu_factors(A)
    assert rel_fro(Lf, L) <= 1e-10 and rel_fro(Uf, U) <= 1e-10
    for trans in (False, True):
        B = oracle.uniform_matrix(n, 3, 9)
        X = B.copy()
        ddla.LUSolve(A, X, trans)
        M = A0.T if trans else A0
        # L U is not well conditioned at n=1000 (unit L with 0.1-scaled entries), so the
        # residual is judged relative to |M| |X|: backward stability of the two solves
        for j in range(X.shape[1]):
            assert backward_error(M, X[:, j], B[:, j]) < 1e-14


def test_lu_nopiv_zero_pivot(ddla):
    A = np.eye(200)
    A[0, 0] = 0.0
    with pytest.raises(ZeroDivisionError):
        ddla.LUDecompose(A)


# ------------------------------------------------------------------------- Cholesky ---
def test_chol_integer_known_answers(ddla):
    c = golden("chol_int.npz")
    for t in range(int(c["count"])):
        A = c[f"A{t}"].copy()
        ddla.CholDecompose(ddla.LOWER, A)
        np.testing.assert_allclose(np.tril(A), c[f"L{t}"], rtol=0, atol=1e-12)
        np.testing.assert_array_equal(np.triu(A, 1), np.triu(c[f"A{t}"], 1))
        U = c[f"A{t}"].copy()
        ddla.CholDecompose(ddla.UPPER, U)
        np.testing.assert_allclose(np.triu(U), c[f"L{t}"].T, rtol=0, atol=1e-12)
        np.testing.assert_array_equal(np.tril(U, -1), np.tril(c[f"A{t}"], -1))


@pytest.mark.parametrize("n", [1, 7, 128, 129, 300, 1000, 2048])
@pytest.mark.parametrize("uplo", [0xC0, 0x60])
def test_chol_matches_oracle(ddla, oracle, n, uplo):
    G = oracle.uniform_matrix(n, n, 77 + n)
    S = G @ G.T + n * np.eye(n)
    So = S.copy()
    assert oracle.potrf(So, uplo) == 0
    A = S.copy()
    ddla.CholDecompose(uplo, A)
    tri = np.tril if uplo == 0xC0 else np.triu
    assert rel_fro(tri(A), tri(So)) <= 1e-13 * n
    B = oracle.uniform_matrix(n, 4, 3)
    X = B.copy()
    ddla.CholeskySolve(uplo, A, X)
    Xo = oracle.potrs(So, B.copy(), uplo)
    assert rel_fro(X, Xo) <= 1e-12


def test_chol_not_spd(ddla):
    S = np.eye(300)
    S[200, 200] = -1.0
    with pytest.raises(ddla.LinalgException):
        ddla.CholDecompose(ddla.LOWER, S.copy())
    d = ddla.DeviceMatrix.from_host(S)
    assert ddla.native.lib().ddla_dpotrf_dev(ddla.LOWER, 300, d.ptr, d.ld) == 201
    d = ddla.DeviceMatrix.from_host(S)
    assert ddla.native.lib().ddla_dpotrf_dev(ddla.UPPER, 300, d.ptr, d.ld) == 201


def test_chol_class_mirror(ddla):
    n = 400
    rng = np.random.default_rng(9)
    G = rng.standard_normal((n, n))
    S = G @ G.T + n * np.eye(n)
    b = rng.uniform(-1, 1, n)
    x = ddla.Cholesky(S).Solve(b)
    assert backward_error(S, x, b) < 1e-14


@pytest.mark.parametrize("n", [1, 31, 32, 33, 64, 100, 128])
def test_cholesky_single_block_matches_unblocked_exactly(ddla, oracle, n):
    # one diagonal block: the blocked-in-LDS kernel performs the unblocked loop's operations
    # (same fmas in the same column order, scaling by the reciprocal), so the factor is
    # bit-identical to the reference's CholeskyLowerUnblock restated in the oracle
    G = oracle.uniform_matrix(n, n, 77 + n)
    S = G @ G.T + n * np.eye(n)
    So = S.copy()
    assert oracle.potrf(So) == 0
    A = S.copy()
    ddla.CholDecompose(ddla.LOWER, A)
    np.testing.assert_array_equal(np.tril(A), np.tril(So))
    np.testing.assert_array_equal(np.triu(A, 1), np.triu(S, 1))


@pytest.mark.slow
def test_lu_8320_panel_with_128_rows_per_workgroup(ddla, oracle):
    # the first panels have more than 8192 rows, so the panel kernel runs with 128 rows per
    # workgroup (the form the n = 16384 factorizations use); the pivot sequence must still be
    # the oracle's
    n = 8320
    A0 = oracle.uniform_matrix(n, n, SEED_C2 + 7)
    Ao = A0.copy()
    info, piv_o = oracle.getrf(Ao)
    assert info == 0
    dA = ddla.DeviceMatrix.from_host(A0)
    piv = ddla.PLUDecompose(dA)
    mism = np.nonzero(piv != piv_o)[0]
    assert mism.size == 0, f"first divergent pivot at step {mism[0] if mism.size else -1}"
    assert rel_fro(dA.to_host(), Ao) <= 1e-13 * n


@pytest.mark.slow
@pytest.mark.parametrize("prec", [2, 3])  # FP16, BF16 trailing updates
def test_mixed_factor_8320_reconstructs(ddla, prec):
    # the mixed LU's FP32 factor through the 128-rows-per-workgroup panel: P A = L U on a
    # sample of rows (the full product would cost the host minutes)
    n = 8320
    A = np.random.default_rng(prec).uniform(-1, 1, (n, n)).astype(np.float32)
    LU, perm, info = ddla.sgetrf_mp(A, prec, ddla.PIVOT_PARTIAL)
    assert info == 0
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert np.abs(np.tril(LU, -1)).max() <= 1.0 + 2.0 ** -22
    rows = np.sort(np.random.default_rng(9).choice(n, 192, replace=False))
    rows = np.unique(np.r_[rows, [0, 1, 127, 128, 8191, 8192, n - 1]])
    Ls = LU[rows].astype(np.float64)
    L = np.where(np.arange(n)[None, :] < rows[:, None], Ls, 0.0)  # row r of L: columns < r
    L[np.arange(len(rows)), rows] = 1.0
    U = np.triu(LU.astype(np.float64))
    PA = A[perm[rows]].astype(np.float64)
    assert np.linalg.norm(PA - L @ U) / np.linalg.norm(PA) <= 5e-6 * np.sqrt(n)


def test_pivots_out_of_range_are_rejected(ddla):
    # a relative pivot that would swap outside the matrix: the reference's SwapRow throws, the
    # library returns DDLA_ERR_INVALID_ARG (host and device forms) and leaves B as it was
    lib = ddla.native.lib()
    n = 50
    A = np.eye(n)
    piv = np.zeros(n, dtype=np.int32)
    piv[45] = 5  # 45 + 5 = 50 is outside
    B = np.ones((n, 2))
    with pytest.raises(ddla.DDLAError) as e:
        ddla.PLUSolve(A, piv, B)
    assert e.value.status == ddla.native.ERR_INVALID_ARG
    dA, dB = ddla.DeviceMatrix.from_host(A), ddla.DeviceMatrix.from_host(B)
    dpiv = ddla._DevInts.from_host(piv)
    st = lib.ddla_dgetrs_dev(0, n, 2, dA.ptr, dA.ld, dpiv.ptr, dB.ptr, dB.ld)
    assert st == ddla.native.ERR_INVALID_ARG
    np.testing.assert_array_equal(dB.to_host(), B)
    st = lib.ddla_dlaswp_dev(ddla.LEFT, 1, n, 2, dB.ptr, dB.ld, dpiv.ptr, n)
    assert st == ddla.native.ERR_INVALID_ARG
    # in range (the last possible swap, and a negative entry, which the reference skips)
    piv[45] = 4
    piv[10] = -3
    dpiv = ddla._DevInts.from_host(piv)
    assert lib.ddla_dlaswp_dev(ddla.LEFT, 1, n, 2, dB.ptr, dB.ld, dpiv.ptr, n) == 0


def test_bounded_waits_give_up_and_recover(ddla, oracle):
    # The panel and the triangular solve are workgroups that wait on each other; every wait is
    # bounded.  A tiny bound set from the host makes the waits give up (every kernel still
    # exits normally): the call reports DDLA_ERR_HIP, and the next calls with the default bound
    # run as usual (no state left behind).
    lib = ddla.native.lib()
    n = 4096
    A0 = oracle.uniform_matrix(n, n, 77)
    d = ddla.DeviceMatrix.from_host(A0)
    dpiv = ddla._DevInts(n)
    assert lib.ddla_set_wait_limit(1) == 0
    try:
        st = lib.ddla_dgetrf_dev(n, d.ptr, d.ld, dpiv.ptr)
    finally:
        assert lib.ddla_set_wait_limit(0) == 0
    assert st == ddla.native.ERR_HIP
    assert "gave up" in ddla.native.last_error()
    # a correct factorization afterwards
    d = ddla.DeviceMatrix.from_host(A0)
    piv = ddla.PLUDecompose(d)
    Ao = A0.copy()
    _, piv_o = oracle.getrf(Ao)
    np.testing.assert_array_equal(piv, piv_o)
    # the solve's waits
    b = oracle.uniform_matrix(n, 1, 78)
    db = ddla.DeviceMatrix.from_host(b)
    dp = ddla._DevInts.from_host(piv)
    assert lib.ddla_set_wait_limit(1) == 0
    try:
        st = lib.ddla_dgetrs_dev(0, n, 1, d.ptr, d.ld, dp.ptr, db.ptr, db.ld)
    finally:
        assert lib.ddla_set_wait_limit(0) == 0
    assert st == ddla.native.ERR_HIP
    db = ddla.DeviceMatrix.from_host(b)
    ddla.PLUSolve(d, piv, db)
    x = db.to_host()
    assert backward_error(A0, x[:, 0], b[:, 0]) < 1e-14
