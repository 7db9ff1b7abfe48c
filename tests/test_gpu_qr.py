"""GPU parity of the blocked Householder QR (UT form) against the CPU oracle's restatement of
QR.cs (itself pinned to LAPACK's dgeqrf): R, the Householder vectors and the T factors within
1e-12 of the largest entry, Q orthogonal, least squares against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mat(m, n, seed):
    return np.random.default_rng(seed).uniform(-1, 1, (m, n))


def _t_upper_parts(T, m, n):
    """The entries of T that carry the factorization: the upper triangle of each block."""
    tb = T.shape[0]
    k = min(m, n)
    mask = np.zeros(T.shape, bool)
    for k0 in range(0, k, tb):
        b = min(tb, k - k0)
        mask[:b, k0:k0 + b] = np.triu(np.ones((b, b), bool))
    return np.where(mask, T, 0.0)


@pytest.mark.parametrize("m,n", [(1, 1), (5, 3), (3, 5), (33, 33), (64, 64), (300, 300),
                                 (700, 520), (100, 300), (600, 257), (1000, 700), (2048, 1024)])
def test_qr_matches_oracle(ddla, oracle, m, n):
    A0 = _mat(m, n, m * 31 + n)
    Ao = A0.copy()
    To = oracle.qr(Ao)
    A = A0.copy()
    T = ddla.QRDecompose(A)
    assert T.shape == To.shape == (ddla.qr_t_rows(m, n), n)
    scale = max(1.0, np.max(np.abs(Ao)))
    assert np.max(np.abs(A - Ao)) <= 1e-12 * scale
    Tm, Tom = _t_upper_parts(T, m, n), _t_upper_parts(To, m, n)
    assert np.max(np.abs(Tm - Tom)) <= 1e-12 * max(1.0, np.max(np.abs(Tom)))


@pytest.mark.parametrize("m,n", [(257, 257), (1500, 600), (400, 900)])
def test_qr_q_orthogonal_and_reconstructs(ddla, m, n):
    A0 = _mat(m, n, m + 7 * n)
    f = ddla.QR(A0)
    Q, R = f.Deconstruct()
    assert np.max(np.abs(Q.T @ Q - np.eye(m))) < 1e-13
    assert np.max(np.abs(Q @ R - A0)) < 1e-12
    with pytest.raises(RuntimeError):
        f.Solve(np.ones(m))
    # Q^T Q B = B through the device entry points
    g = ddla.QR(A0)
    B = _mat(m, 5, 3)
    Bq = g.ApplyQT(B.copy())
    np.testing.assert_allclose(g.ApplyQ(Bq), B, rtol=0, atol=1e-12)


@pytest.mark.parametrize("m,n,nrhs", [(40, 40, 1), (500, 300, 3), (2000, 1000, 1), (1000, 1000, 16)])
def test_qr_least_squares_matches_oracle(ddla, oracle, m, n, nrhs):
    rng = np.random.default_rng(m + n + nrhs)
    A0 = rng.uniform(-1, 1, (m, n))
    B = rng.uniform(-1, 1, (m, nrhs))
    Ao = A0.copy()
    To = oracle.qr(Ao)
    Xo = oracle.qr_solve(Ao, To, B)
    dA = ddla.DeviceMatrix.from_host(A0)
    dT = ddla.DeviceMatrix.from_host(np.zeros((ddla.qr_t_rows(m, n), n)))
    ddla.QRDecompose(dA, dT)
    dB = ddla.DeviceMatrix.from_host(B)
    dX = ddla.DeviceMatrix.from_host(np.zeros((n, nrhs)))
    ddla.QRSolve(dA, dT, dB, dX)
    X = dX.to_host()
    np.testing.assert_array_equal(dB.to_host(), B)  # B is kept
    assert np.max(np.abs(X - Xo)) <= 1e-10 * max(1.0, np.max(np.abs(Xo)))
    # normal equations: A^T (A X - B) = 0
    r = A0.T @ (A0 @ X - B)
    assert np.max(np.abs(r)) < 1e-10 * np.max(np.abs(A0.T @ B))


def test_qr_host_strided_views(ddla, oracle):
    m, n = 530, 300
    A0 = _mat(m, n, 11)
    Ao = A0.copy()
    To = oracle.qr(Ao)
    Af = np.asfortranarray(A0.copy())
    big = np.zeros((ddla.qr_t_rows(m, n), 2 * n))
    Tv = big[:, ::2]
    ddla.QRDecompose(Af, Tv)
    assert np.max(np.abs(Af - Ao)) <= 1e-12
    assert np.max(np.abs(_t_upper_parts(Tv, m, n) - _t_upper_parts(To, m, n))) <= 1e-12 * np.max(np.abs(To))
    np.testing.assert_array_equal(big[:, 1::2], 0.0)
    b = _mat(m, 1, 12)[:, 0]
    x = ddla.QRSolve(Af, Tv, b)
    np.testing.assert_allclose(x, np.linalg.lstsq(A0, b, rcond=None)[0], rtol=0, atol=1e-11)


def test_qr_limits_and_bounded_waits(ddla, oracle):
    lib = ddla.native.lib()
    d = ddla.DeviceMatrix.from_host(np.zeros((4, 4)))
    assert lib.ddla_dgeqrf_dev(32769, 1, d.ptr, 1, d.ptr, 1) == ddla.native.ERR_UNSUPPORTED
    assert "32768" in ddla.native.last_error()
    assert lib.ddla_dqr_solve_dev(3, 4, 1, d.ptr, 4, d.ptr, 4, d.ptr, 1, d.ptr, 1) == \
        ddla.native.ERR_INVALID_ARG
    # a tiny poll bound set from the host: the panel's waits give up (every kernel exits), the
    # call reports it, and the next factorization is correct
    m, n = 8192, 64
    A0 = _mat(m, n, 99)
    dA = ddla.DeviceMatrix.from_host(A0)
    dT = ddla.DeviceMatrix.from_host(np.zeros((ddla.qr_t_rows(m, n), n)))
    assert lib.ddla_set_wait_limit(1) == 0
    try:
        st = lib.ddla_dgeqrf_dev(m, n, dA.ptr, dA.ld, dT.ptr, dT.ld)
    finally:
        assert lib.ddla_set_wait_limit(0) == 0
    assert st == ddla.native.ERR_HIP
    assert "gave up" in ddla.native.last_error()
    A = A0.copy()
    ddla.QRDecompose(A)
    Ao = A0.copy()
    oracle.qr(Ao)
    assert np.max(np.abs(A - Ao)) <= 1e-12
