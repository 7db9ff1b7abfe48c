"""GPU parity of the level-3 BLAS entry points with BLIS argument lists (ddla_dgemmt, ddla_dsyrk,
ddla_dsyr2k, ddla_dsymm, ddla_dtrmm, ddla_dtrmm3, ddla_dtrsm and their strided device forms)
against the CPU oracle, on row-major, column-major and colStride = 4 views as the reference's
level-3 tests use them (Tests/TestManagedBlas/Level3m/TestTrMM.cs:12-38, TestGeMM.cs:13-38).
Tolerance: relative Frobenius error <= 1e-13 * k (k the inner dimension)."""
import ctypes

import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu

LO, UP = 0xC0, 0x60
N, T = 0x00, 0x08


def views(M):
    """Row-major, column-major and colStride = 4 views holding the same values."""
    m, n = M.shape
    big = np.zeros((m, 4 * n))
    big[:, ::4] = M
    return {"row": M.copy(), "col": np.asfortranarray(M), "stride4": big[:, ::4]}


def tri_mask(m, uplo):
    return np.tril(np.ones((m, m), bool)) if uplo == LO else np.triu(np.ones((m, m), bool))


@pytest.mark.parametrize("layout", ["row", "col", "stride4"])
@pytest.mark.parametrize("uplo", [LO, UP])
@pytest.mark.parametrize("ta,tb", [(N, N), (T, N), (N, T), (T, T)])
def test_gemmt_host(ddla, oracle, layout, uplo, ta, tb):
    m, k = 300, 145
    A = oracle.uniform_matrix(*(k, m) if ta else (m, k), 21)
    B = oracle.uniform_matrix(*(m, k) if tb else (k, m), 22)
    C0 = oracle.uniform_matrix(m, m, 23)
    ref = oracle.gemmt(uplo, ta, tb, -0.5, A, B, 1.5, C0.copy())
    Av, Bv, Cv = views(A)[layout], views(B)[layout], views(C0)[layout]
    ddla.BlasProvider.GeMMT(uplo, ta, tb, -0.5, Av, Bv, 1.5, Cv)
    C = np.array(Cv)
    mask = tri_mask(m, uplo)
    assert rel_fro(C[mask], ref[mask]) <= 1e-13 * k
    np.testing.assert_array_equal(C[~mask], C0[~mask])


@pytest.mark.parametrize("uplo", [LO, UP])
def test_gemmt_device_and_edges(ddla, oracle, uplo):
    m, k = 1000, 128  # enough tiles for the 128 x 128 kernel
    A = oracle.uniform_matrix(m, k, 24)
    B = oracle.uniform_matrix(k, m, 25)
    C0 = oracle.uniform_matrix(m, m, 26)
    ref = oracle.gemmt(uplo, N, N, 1.0, A, B, 0.0, C0.copy())
    dA, dB, dC = (ddla.DeviceMatrix.from_host(X) for X in (A, B, C0))
    ddla.BlasProvider.GeMMT(uplo, N, N, 1.0, dA, dB, 0.0, dC)
    C = dC.to_host()
    mask = tri_mask(m, uplo)
    assert rel_fro(C[mask], ref[mask]) <= 1e-13 * k
    np.testing.assert_array_equal(C[~mask], C0[~mask])
    # k == 0 only scales the triangle; beta == 0 clears NaN there
    Cn = np.full((5, 5), np.nan)
    ddla.BlasProvider.GeMMT(uplo, N, N, 1.0, np.zeros((5, 0)), np.zeros((0, 5)), 0.0, Cn)
    assert np.all(Cn[tri_mask(5, uplo)] == 0.0) and np.all(np.isnan(Cn[~tri_mask(5, uplo)]))


@pytest.mark.parametrize("layout", ["row", "col", "stride4"])
@pytest.mark.parametrize("uplo", [LO, UP])
@pytest.mark.parametrize("trans", [N, T])
def test_syrk_host(ddla, oracle, layout, uplo, trans):
    m, k = 257, 130
    A = oracle.uniform_matrix(*(k, m) if trans else (m, k), 31)
    C0 = oracle.uniform_matrix(m, m, 32)
    ref = oracle.syrk(uplo, trans, 0.8, A, 1.1, C0.copy())
    Av, Cv = views(A)[layout], views(C0)[layout]
    ddla.BlasProvider.SyRk(uplo, trans, 0.8, Av, 1.1, Cv)
    C = np.array(Cv)
    mask = tri_mask(m, uplo)
    assert rel_fro(C[mask], ref[mask]) <= 1e-13 * k
    np.testing.assert_array_equal(C[~mask], C0[~mask])


@pytest.mark.parametrize("layout", ["row", "col", "stride4"])
@pytest.mark.parametrize("uplo", [LO, UP])
@pytest.mark.parametrize("ta,tb", [(N, N), (T, T), (N, T)])
def test_syr2k(ddla, oracle, layout, uplo, ta, tb):
    m, k = 200, 77
    A = oracle.uniform_matrix(*(k, m) if ta else (m, k), 33)
    B = oracle.uniform_matrix(*(k, m) if tb else (m, k), 34)
    C0 = oracle.uniform_matrix(m, m, 35)
    ref = oracle.syr2k(uplo, ta, tb, 1.3, A, B, 0.6, C0.copy())
    Av, Bv, Cv = views(A)[layout], views(B)[layout], views(C0)[layout]
    ddla.BlasProvider.SyR2k(uplo, ta, tb, 1.3, Av, Bv, 0.6, Cv)
    C = np.array(Cv)
    mask = tri_mask(m, uplo)
    assert rel_fro(C[mask], ref[mask]) <= 1e-13 * k
    np.testing.assert_array_equal(C[~mask], C0[~mask])
    # the device form gives the same result
    dA, dB, dC = (ddla.DeviceMatrix.from_host(X) for X in (A, B, C0))
    ddla.BlasProvider.SyR2k(uplo, ta, tb, 1.3, dA, dB, 0.6, dC)
    assert rel_fro(dC.to_host()[mask], ref[mask]) <= 1e-13 * k


@pytest.mark.parametrize("layout", ["row", "col", "stride4"])
@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("uplo", [LO, UP])
@pytest.mark.parametrize("m,n", [(64, 64), (768, 64), (64, 768)])
def test_symm(ddla, oracle, layout, side, uplo, m, n):
    # shapes and scalars of TestSyMM.cs:39-66
    k = m if side == 0 else n
    A = oracle.uniform_matrix(k, k, 41)
    B = oracle.uniform_matrix(m, n, 42)
    C0 = oracle.uniform_matrix(m, n, 43)
    ref = oracle.symm(side, uplo, N, 1 / np.pi, A, B, np.e, C0.copy())
    Av, Bv, Cv = views(A)[layout], views(B)[layout], views(C0)[layout]
    ddla.BlasProvider.SyMM(side, uplo, N, 1 / np.pi, Av, Bv, np.e, Cv)
    assert rel_fro(np.array(Cv), ref) <= 1e-13 * k


def test_symm_reads_only_its_triangle(ddla, oracle):
    m, n = 100, 60
    A = oracle.uniform_matrix(m, m, 44)
    A_poisoned = A.copy()
    A_poisoned[np.triu_indices(m, 1)] = np.nan  # lower triangle stored
    B = oracle.uniform_matrix(m, n, 45)
    C = np.zeros((m, n))
    ddla.BlasProvider.SyMM(0, LO, T, 1.0, A_poisoned, np.ascontiguousarray(B.T), 0.0, C)
    ref = oracle.symm(0, LO, N, 1.0, A, B, 0.0, np.zeros((m, n)))
    assert rel_fro(C, ref) <= 1e-13 * m


@pytest.mark.parametrize("layout", ["row", "col", "stride4"])
@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("uplo", [LO, UP])
@pytest.mark.parametrize("trans", [N, T])
@pytest.mark.parametrize("diag", [0x000, 0x100])
def test_trmm(ddla, oracle, layout, side, uplo, trans, diag):
    m, n = 300, 130
    k = m if side == 0 else n
    A = oracle.uniform_matrix(k, k, 51)
    A[~tri_mask(k, uplo)] = np.nan  # outside the triangle is never read
    if diag == 0x100:
        A[np.diag_indices(k)] = np.nan
    B0 = oracle.uniform_matrix(m, n, 52)
    ref = oracle.trmm(side, uplo, trans, diag, 1.2, np.nan_to_num(A), B0.copy())
    Av, Bv = views(A)[layout], views(B0)[layout]
    ddla.BlasProvider.TrMM(side, uplo, trans, diag, 1.2, Av, Bv)
    assert rel_fro(np.array(Bv), ref) <= 1e-13 * k


@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("uplo", [LO, UP])
@pytest.mark.parametrize("ta,tb", [(N, N), (T, T)])
def test_trmm3_and_device_forms(ddla, oracle, side, uplo, ta, tb):
    m, n = 260, 190
    k = m if side == 0 else n
    A = oracle.uniform_matrix(k, k, 61)
    B = oracle.uniform_matrix(*(n, m) if tb else (m, n), 62)
    C0 = oracle.uniform_matrix(m, n, 63)
    ref = oracle.trmm3(side, uplo, ta, 0x000, tb, 0.9, A, B, -1.0, C0.copy())
    C = np.asfortranarray(C0)
    ddla.BlasProvider.TrMM3(side, uplo, ta, 0x000, tb, 0.9, A, B, -1.0, C)
    assert rel_fro(C, ref) <= 1e-13 * k
    dA, dB, dC = (ddla.DeviceMatrix.from_host(X) for X in (A, B, C0))
    ddla.BlasProvider.TrMM3(side, uplo, ta, 0x000, tb, 0.9, dA, dB, -1.0, dC)
    assert rel_fro(dC.to_host(), ref) <= 1e-13 * k
    # device TRMM over B in place
    Bsq = oracle.uniform_matrix(m, n, 64)
    refb = oracle.trmm(side, uplo, ta, 0x100, -2.0, A, Bsq.copy())
    dB2 = ddla.DeviceMatrix.from_host(Bsq)
    ddla.BlasProvider.TrMM(side, uplo, ta, 0x100, -2.0, dA, dB2)
    assert rel_fro(dB2.to_host(), refb) <= 1e-13 * k


@pytest.mark.parametrize("layout", ["row", "col", "stride4"])
@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("uplo", [LO, UP])
@pytest.mark.parametrize("trans", [N, T])
@pytest.mark.parametrize("diag", [0x000, 0x100])
def test_trsm_host_blis_form(ddla, oracle, layout, side, uplo, trans, diag):
    m, n = 300, 70
    k = m if side == 0 else n
    A = oracle.uniform_matrix(k, k, 71) * 0.05 + 2 * np.eye(k)
    B = oracle.uniform_matrix(m, n, 72)
    ref = oracle.trsm(side, uplo, trans, diag, 1.5, A, B.copy())
    Av, Bv = views(A)[layout], views(B)[layout]
    ddla.BlasProvider.TrSM(side, uplo, trans, diag, 1.5, Av, Bv)
    assert rel_fro(np.array(Bv), ref) <= 1e-13 * k


def test_level3_bad_arguments(ddla):
    lib = ddla.native.lib()
    one = ctypes.c_double(1.0)
    st = lib.ddla_dgemmt(0x12, N, N, 2, 2, ctypes.byref(one), None, 1, 1, None, 1, 1,
                         ctypes.byref(one), None, 1, 1)
    assert st == ddla.native.ERR_INVALID_ARG
    st = lib.ddla_dsymm(0, LO, 0x20, N, 2, 2, ctypes.byref(one), None, 1, 1, None, 1, 1,
                        ctypes.byref(one), None, 1, 1)
    assert st == ddla.native.ERR_INVALID_ARG
    st = lib.ddla_dtrmm(0, LO, N, 0x3, 2, 2, ctypes.byref(one), None, 1, 1, None, 1, 1)
    assert st == ddla.native.ERR_INVALID_ARG
    with pytest.raises(ValueError):
        ddla.BlasProvider.TrMM(0, LO, N, 0, 1.0, np.eye(3), np.zeros((4, 2)))
    with pytest.raises(ValueError):
        ddla.BlasProvider.GeMMT(LO, N, N, 1.0, np.ones((3, 2)), np.ones((3, 3)), 0.0,
                                np.zeros((3, 3)))
