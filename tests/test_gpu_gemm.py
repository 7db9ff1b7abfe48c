"""GPU parity of the BLAS pieces (FP64 MFMA GEMM, SYRK, TRSM, AMAX, device fill) against
the CPU oracle.  Tolerance for GEMM: relative Frobenius error <= 1e-13 * K."""
import ctypes

import numpy as np
import pytest

from conftest import golden, rel_fro

pytestmark = pytest.mark.gpu

SEED_C1 = 0xDD1A0001


def test_device_fill_matches_oracle_stream(ddla, oracle):
    d = ddla.DeviceMatrix.random(37, 53, 12345)
    np.testing.assert_array_equal(d.to_host(), oracle.uniform_matrix(37, 53, 12345))


def test_gemm_1024_parity(ddla, oracle):
    # config: M=N=K=1024, alpha=1, beta=0, row-major, uniform[-1,1] from a fixed seed
    n = 1024
    A = oracle.uniform_matrix(n, n, SEED_C1)
    B = oracle.uniform_matrix(n, n, SEED_C1 + 1)
    ref = oracle.matmul(A, B)
    C = np.full((n, n), np.nan)
    ddla.BlasProvider.GeMM(1.0, A, B, 0.0, C)
    assert rel_fro(C, ref) <= 1e-13 * n
    # device-resident operands, same result bit for bit
    dA, dB = ddla.DeviceMatrix.from_host(A), ddla.DeviceMatrix.from_host(B)
    dC = ddla.DeviceMatrix(n, n)
    ddla.BlasProvider.GeMM(1.0, dA, dB, 0.0, dC)
    np.testing.assert_array_equal(dC.to_host(), C)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 33, 9), (130, 70, 300), (257, 129, 145),
                                   (64, 64, 64), (1000, 1000, 37), (300, 2000, 513)])
def test_gemm_shapes(ddla, oracle, m, n, k):
    A = oracle.uniform_matrix(m, k, m + 1)
    B = oracle.uniform_matrix(k, n, n + 2)
    C0 = oracle.uniform_matrix(m, n, k + 3)
    ref = oracle.gemm(2.0, A, B, 1.2, C0.copy())
    C = C0.copy()
    ddla.BlasProvider.GeMM(2.0, A, B, 1.2, C)
    assert rel_fro(C, ref) <= 1e-13 * k


@pytest.mark.parametrize("k", [1, 17, 128])
def test_rank_k_update_large_tiles(ddla, oracle, k):
    # C -= A B with beta = 1 at the factorizations' update shape (K one block or less), large
    # enough for the 128 x 128 tiles: interior tiles fetch C during the last K step, the
    # ragged right and bottom edges take the bounds-checked store
    m, n = 3000, 2950  # 24 x 24 tiles of 128: at least two per CU, so the 128 x 128 kernel runs
    A = oracle.uniform_matrix(m, k, 41 + k)
    B = oracle.uniform_matrix(k, n, 42 + k)
    C0 = oracle.uniform_matrix(m, n, 43 + k)
    ref = oracle.gemm(-1.0, A, B, 1.0, C0.copy())
    dA, dB, dC = (ddla.DeviceMatrix.from_host(X) for X in (A, B, C0))
    ddla.BlasProvider.GeMM(-1.0, dA, dB, 1.0, dC)
    assert rel_fro(dC.to_host(), ref) <= 1e-13 * max(k, 8)


@pytest.mark.parametrize("uplo", [0xC0, 0x60])
def test_syrk_large_tiles(ddla, oracle, uplo):
    # SYRK update at 128 x 128 tiles: tiles wholly inside the stored triangle take the early C
    # fetch, diagonal tiles the masked store; the other triangle stays untouched
    n, k = 3000, 128
    A = oracle.uniform_matrix(n, k, 51)
    C0 = oracle.uniform_matrix(n, n, 52)
    ref = oracle.syrk(uplo, 0x00, -1.0, A, 1.0, C0.copy())
    dA, dC = ddla.DeviceMatrix.from_host(A), ddla.DeviceMatrix.from_host(C0)
    ddla.BlasProvider.SyRk(uplo, 0x00, -1.0, dA, 1.0, dC)
    C = dC.to_host()
    mask = np.tril(np.ones((n, n), bool)) if uplo == 0xC0 else np.triu(np.ones((n, n), bool))
    assert rel_fro(C[mask], ref[mask]) <= 1e-13 * k
    np.testing.assert_array_equal(C[~mask], C0[~mask])


@pytest.mark.parametrize("ta", [ddla_t for ddla_t in (0x00, 0x08)])
@pytest.mark.parametrize("tb", [0x00, 0x08])
def test_gemm_transposes(ddla, oracle, ta, tb):
    m, n, k = 150, 90, 200
    A = oracle.uniform_matrix(*(k, m) if ta else (m, k), 1)
    B = oracle.uniform_matrix(*(n, k) if tb else (k, n), 2)
    C0 = oracle.uniform_matrix(m, n, 3)
    ref = oracle.gemm(-1.0, A, B, 1.0, C0.copy(), ta, tb)
    C = C0.copy()
    ddla.BlasProvider.GeMM(-1.0, A, B, 1.0, C, ta, tb)
    assert rel_fro(C, ref) <= 1e-13 * k


def test_gemm_strided_and_colmajor_views(ddla, oracle):
    # the reference tests row-major, column-major and colStride=4 views (TestGeMM.cs:13-38)
    m, n, k = 70, 45, 160
    big = oracle.uniform(4 * m * k + 16, 9)
    A = np.lib.stride_tricks.as_strided(big, shape=(m, k), strides=(8 * 4 * k, 8 * 4))
    B = np.asfortranarray(oracle.uniform_matrix(k, n, 10))
    C = np.asfortranarray(oracle.uniform_matrix(m, n, 11))
    ref = oracle.gemm(2.0, np.ascontiguousarray(A), np.ascontiguousarray(B), 1.2,
                      np.ascontiguousarray(C))
    ddla.BlasProvider.GeMM(2.0, A, B, 1.2, C)
    assert rel_fro(C, ref) <= 1e-13 * k
    # general strides on the output too
    Cs = np.zeros((m, 3 * n))[:, ::3]
    ddla.BlasProvider.GeMM(1.0, A, B, 0.0, Cs)
    assert rel_fro(Cs, oracle.matmul(np.ascontiguousarray(A), np.ascontiguousarray(B))) <= 1e-13 * k


def test_gemm_integer_known_answers(ddla):
    g = golden("gemm_int.npz")
    for t in range(int(g["count"])):
        A, B = g[f"A{t}"], g[f"B{t}"]
        C = np.zeros((A.shape[0], B.shape[1]))
        ddla.BlasProvider.GeMM(1.0, A, B, 0.0, C)
        np.testing.assert_array_equal(C, g[f"C{t}"])


def test_gemm_beta_zero_clears_nan_and_k_zero(ddla):
    A = np.zeros((5, 0))
    B = np.zeros((0, 6))
    C = np.full((5, 6), np.nan)
    ddla.BlasProvider.GeMM(1.0, A, B, 0.0, C)
    assert np.all(C == 0.0)
    C = np.ones((5, 6))
    ddla.BlasProvider.GeMM(1.0, A, B, 3.0, C)
    assert np.all(C == 3.0)
    A = np.ones((4, 4))
    C = np.full((4, 4), np.nan)
    ddla.BlasProvider.GeMM(1.0, A, A, 0.0, C)
    assert np.all(C == 4.0)


def test_gemm_bad_arguments_raise(ddla):
    with pytest.raises(ValueError):
        ddla.BlasProvider.GeMM(1.0, np.ones((3, 4)), np.ones((5, 2)), 0.0, np.zeros((3, 2)))
    lib = ddla.native.lib()
    al = ctypes.c_double(1.0)
    st = lib.ddla_dgemm(0x33, 0, 1, 1, 1, ctypes.byref(al), None, 1, 1, None, 1, 1,
                        ctypes.byref(al), None, 1, 1)
    assert st == ddla.native.ERR_INVALID_ARG


@pytest.mark.parametrize("uplo", [0xC0, 0x60])
@pytest.mark.parametrize("trans", [0x00, 0x08])
def test_syrk(ddla, oracle, uplo, trans):
    n, k = 300, 128
    A = oracle.uniform_matrix(*(k, n) if trans else (n, k), 9)
    C0 = oracle.uniform_matrix(n, n, 10)
    ref = oracle.syrk(uplo, trans, -1.0, A, 1.0, C0.copy())
    dA, dC = ddla.DeviceMatrix.from_host(A), ddla.DeviceMatrix.from_host(C0)
    ddla.BlasProvider.SyRk(uplo, trans, -1.0, dA, 1.0, dC)
    C = dC.to_host()
    mask = np.tril(np.ones((n, n), bool)) if uplo == 0xC0 else np.triu(np.ones((n, n), bool))
    assert rel_fro(C[mask], ref[mask]) <= 1e-13 * k
    np.testing.assert_array_equal(C[~mask], C0[~mask])


@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("uplo", [0xC0, 0x60])
@pytest.mark.parametrize("trans", [0x00, 0x08])
@pytest.mark.parametrize("diag", [0x000, 0x100])
def test_trsm_variants(ddla, oracle, side, uplo, trans, diag):
    m, n = 300, 70
    ka = m if side == 0 else n
    A = oracle.uniform_matrix(ka, ka, 5) * 0.05 + 2 * np.eye(ka)
    B = oracle.uniform_matrix(m, n, 6)
    ref = oracle.trsm(side, uplo, trans, diag, 1.5, A, B.copy())
    dA, dB = ddla.DeviceMatrix.from_host(A), ddla.DeviceMatrix.from_host(B)
    ddla.BlasProvider.TrSM(side, uplo, trans, diag, 1.5, dA, dB)
    assert rel_fro(dB.to_host(), ref) <= 1e-13 * ka


def test_amax(ddla):
    x = np.array([1.0, -3.0, 3.0, 2.0, np.nan, -3.0])
    d = ddla.DeviceMatrix.from_host(x.reshape(-1, 1))
    assert ddla.BlasProvider.AMax(d) == 1
    d = ddla.DeviceMatrix.from_host(np.array([[np.nan], [0.5], [-0.5]]))
    assert ddla.BlasProvider.AMax(d) == 1
    rng = np.random.default_rng(3)
    v = rng.standard_normal(100000)
    v[77777] = 50.0
    v[88888] = -50.0
    d = ddla.DeviceMatrix.from_host(v.reshape(-1, 1))
    assert ddla.BlasProvider.AMax(d) == 77777
