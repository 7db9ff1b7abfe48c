"""Mixed-precision LU / Cholesky + iterative refinement: FP64-level backward error and
forward error against the FP64 oracle solution, on matrices of known condition number."""
import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu


def conditioned(n, kappa, seed, spd=False):
    """A = U diag(s) V^T with U, V from QR of Gaussians and s geometric from 1 to 1/kappa
    (the construction of the mixed-precision benchmark at small n)."""
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = kappa ** (-np.arange(n) / (n - 1))
    if spd:
        return (U * s) @ U.T
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (U * s) @ V.T


def run(ddla, fn, A, nrhs, prec, seed=1, tol=1e-14, terms=0, **kw):
    n = A.shape[0]
    B = np.random.default_rng(seed).uniform(-1, 1, (n, nrhs))
    dA, dB = ddla.DeviceMatrix.from_host(A), ddla.DeviceMatrix.from_host(B)
    dX = ddla.DeviceMatrix(n, nrhs)
    info = fn(dA, dB, dX, prec=prec, tol=tol, terms=terms, **kw)
    np.testing.assert_array_equal(dA.to_host(), A)  # A is not modified
    return info, dX.to_host(), B


@pytest.mark.parametrize("prec", [1, 2, 3])  # FP32, FP16, BF16
@pytest.mark.parametrize("kappa", [1e2, 1e4])
@pytest.mark.parametrize("nrhs", [1, 5])
def test_gesv_mpir_converges(ddla, oracle, prec, kappa, nrhs):
    n = 1024
    A = conditioned(n, kappa, 11)
    info, X, B = run(ddla, ddla.gesv_mpir, A, nrhs, prec)
    assert info.status == 0, (info.status, info.iters, info.backward_error)
    assert info.backward_error <= 1e-14
    for j in range(nrhs):
        r = B[:, j] - A @ X[:, j]
        be = np.linalg.norm(r, np.inf) / (np.linalg.norm(A, np.inf) * np.linalg.norm(X[:, j], np.inf))
        assert be <= 2e-14
    # forward error against the FP64 oracle solution: both are backward stable to ~1e-14,
    # so they may differ by up to about kappa_inf(A) * 1e-14 (kappa_inf <= sqrt(n) kappa_2)
    Ao = A.copy()
    _, piv = oracle.getrf(Ao)
    Xo = oracle.getrs(Ao, piv, B.copy())
    assert rel_fro(X, Xo) <= kappa * np.sqrt(n) * 1e-14


@pytest.mark.parametrize("prec", [1, 2, 3])
def test_posv_mpir_converges(ddla, oracle, prec):
    n = 1024
    A = conditioned(n, 1e4, 12, spd=True)
    A = (A + A.T) / 2
    info, X, B = run(ddla, ddla.posv_mpir, A, 2, prec)
    assert info.status == 0, (info.status, info.iters, info.backward_error)
    assert info.backward_error <= 1e-14
    So = A.copy()
    assert oracle.potrf(So) == 0
    Xo = oracle.potrs(So, B.copy())
    assert rel_fro(X, Xo) <= 1e4 * 1e-14


def test_posv_mpir_fallback_to_fp32(ddla):
    # lambda down to 1e-5: a single-term BF16 factor (u = 2^-8) of this SPD matrix loses
    # positive definiteness, an FP32 factor does not
    n = 512
    A = conditioned(n, 1e5, 13, spd=True)
    A = (A + A.T) / 2
    info, X, B = run(ddla, ddla.posv_mpir, A, 1, 3, terms=1)
    assert info.fallback == 1
    assert info.status == 0 and info.backward_error <= 1e-14


@pytest.mark.parametrize("prec", [1, 2, 3])
@pytest.mark.parametrize("n", [1021, 1024])  # 1021: butterflies on a padded order
@pytest.mark.parametrize("nrhs", [1, 3])
def test_gesv_mpir_butterfly(ddla, oracle, prec, n, nrhs):
    kappa = 1e4
    A = conditioned(n, kappa, 21)
    info, X, B = run(ddla, ddla.gesv_mpir, A, nrhs, prec, pivoting=ddla.PIVOT_BUTTERFLY)
    assert info.status == 0, (info.status, info.iters, info.backward_error, info.fallback)
    # either the pivot-free factor refined on its own or the solver switched to partial
    # pivoting (element growth of the transformed elimination can exceed what an FP32 /
    # FP16 / BF16 factor refines from); the answer reaches FP64 accuracy in both cases
    assert info.fallback in (0, 2), (info.iters, info.backward_error)
    assert info.backward_error <= 1e-14
    Ao = A.copy()
    _, piv = oracle.getrf(Ao)
    Xo = oracle.getrs(Ao, piv, B.copy())
    assert rel_fro(X, Xo) <= kappa * np.sqrt(n) * 1e-14


def test_gesv_mpir_butterfly_permutation(ddla):
    # a permutation matrix: LU without pivoting breaks down on it at once; after the
    # two-sided butterfly transform it must not (or the solver falls back to partial
    # pivoting) -- either way the solve reaches FP64 accuracy
    n = 512
    P = np.eye(n)[np.random.default_rng(4).permutation(n)]
    info, X, B = run(ddla, ddla.gesv_mpir, P, 2, 2, pivoting=ddla.PIVOT_BUTTERFLY)
    assert info.status == 0 and info.backward_error <= 1e-14
    np.testing.assert_allclose(P @ X, B, rtol=0, atol=1e-13)


def test_gesv_mpir_rejects_singular(ddla):
    n = 256
    A = np.random.default_rng(3).uniform(-1, 1, (n, n))
    A[:, 100] = 0.0
    info, _, _ = run(ddla, ddla.gesv_mpir, A, 1, 2)
    assert info.status == 101


@pytest.mark.parametrize("fn", ["gesv", "posv"])
def test_mpir_many_rhs(ddla, oracle, fn):
    # more right-hand sides than one persistent solve takes (chunks of 8, 8 and 4)
    n, nrhs, kappa = 1024, 20, 1e4
    spd = fn == "posv"
    A = conditioned(n, kappa, 41, spd=spd)
    if spd:
        A = (A + A.T) / 2
    info, X, B = run(ddla, ddla.gesv_mpir if fn == "gesv" else ddla.posv_mpir, A, nrhs, 2)
    assert info.status == 0 and info.backward_error <= 1e-14
    for j in range(nrhs):
        r = B[:, j] - A @ X[:, j]
        be = np.linalg.norm(r, np.inf) / (np.linalg.norm(A, np.inf) * np.linalg.norm(X[:, j], np.inf))
        assert be <= 2e-14, j


@pytest.mark.parametrize("prec", [1, 2, 3])
@pytest.mark.parametrize("n", [129, 257, 300, 1021, 1300])
def test_partial_factor_reconstructs(ddla, prec, n):
    """The mixed LU's FP32 factor with partial pivoting (label-tracking panel, the next block's
    interchanges fused into its unit-lower solve, FP32 look-ahead update, FP16 / BF16 / FP32 far
    updates) reproduces P A, and its multipliers are bounded by 1 (partial pivoting; l is the
    entry times the rounded reciprocal of the pivot, so up to one rounding above 1)."""
    A = np.random.default_rng(7 * n + prec).uniform(-1, 1, (n, n)).astype(np.float32)
    LU, perm, info = ddla.sgetrf_mp(A, prec, ddla.PIVOT_PARTIAL)
    assert info == 0
    assert sorted(perm.tolist()) == list(range(n))
    L = np.tril(LU.astype(np.float64), -1) + np.eye(n)
    U = np.triu(LU.astype(np.float64))
    PA = A.astype(np.float64)[perm]
    assert np.linalg.norm(PA - L @ U) / np.linalg.norm(PA) <= 5e-6 * np.sqrt(n)
    assert np.abs(np.tril(LU, -1)).max() <= 1.0 + 2.0 ** -22
