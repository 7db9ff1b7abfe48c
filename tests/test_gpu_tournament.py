"""Tournament pivoting (the mixed LU's CALU-style panel) on the GPU against the CPU
restatement in the oracle: the first panel's selected rows and factored L11 \\ U11 bit for bit,
the whole FP32 factor by reconstruction, and the refinement built on it."""
import numpy as np
import pytest

from conftest import rel_fro
from test_gpu_mixed import conditioned, run

pytestmark = pytest.mark.gpu


def reconstruct_err(A, LU, perm):
    n = A.shape[0]
    L = np.tril(LU.astype(np.float64), -1) + np.eye(n)
    U = np.triu(LU.astype(np.float64))
    PA = A.astype(np.float64)[perm]
    return np.linalg.norm(PA - L @ U) / np.linalg.norm(PA)


@pytest.mark.parametrize("n", [128, 256, 1300, 2100])  # 1300 / 2100: trees with odd levels
def test_first_panel_matches_oracle_bitwise(ddla, oracle, n):
    A = np.random.default_rng(n).uniform(-1, 1, (n, n)).astype(np.float32)
    LU, perm, info = ddla.sgetrf_mp(A, ddla.PREC_FP32, ddla.PIVOT_TOURNAMENT)
    assert info == 0
    zero, sel, top, _ = oracle.tslu(A[:, :128])
    assert zero == 0
    np.testing.assert_array_equal(perm[:128], sel)
    np.testing.assert_array_equal(LU[:128, :128], top)
    assert sorted(perm.tolist()) == list(range(n))
    assert reconstruct_err(A, LU, perm) <= 2e-6 * np.sqrt(n)


@pytest.mark.parametrize("n", [700, 2100])
def test_designed_pivots_through_the_tree(ddla, oracle, n):
    # unit rows scattered over all stacks (the rest zero): every merge passes them on and
    # the panel's pivots are the unit rows in column order
    A = np.random.default_rng(n + 1).uniform(-1, 1, (n, n)).astype(np.float32)
    P, where = _designed(n, n)
    A[:, :128] = P
    LU, perm, info = ddla.sgetrf_mp(A, ddla.PREC_FP32, ddla.PIVOT_TOURNAMENT)
    np.testing.assert_array_equal(perm[:128], where)
    zero, sel, top, _ = oracle.tslu(P)
    np.testing.assert_array_equal(LU[:128, :128], top)


def _designed(m, seed):
    rng = np.random.default_rng(seed)
    P = np.zeros((m, 128), dtype=np.float32)
    where = rng.permutation(m)[:128]
    for j, r in enumerate(where):
        P[r, j] = 1.0
    return P, where


@pytest.mark.parametrize("prec", [1, 2, 3])
@pytest.mark.parametrize("n", [1000, 1300])
def test_tournament_factor_reconstructs(ddla, prec, n):
    # FP16 / BF16 trailing updates use split operands (FP32-level products)
    A = np.random.default_rng(3 * n + prec).uniform(-1, 1, (n, n)).astype(np.float32)
    LU, perm, info = ddla.sgetrf_mp(A, prec, ddla.PIVOT_TOURNAMENT)
    assert info == 0
    assert reconstruct_err(A, LU, perm) <= 5e-6 * np.sqrt(n)
    # growth: multipliers stay moderate (partial pivoting would bound them by 1)
    assert np.abs(np.tril(LU, -1)).max() <= 8.0


@pytest.mark.parametrize("prec", [1, 2, 3])
@pytest.mark.parametrize("n", [1021, 2100])  # 1021: a last panel narrower than 128
@pytest.mark.parametrize("nrhs", [1, 3])
def test_gesv_mpir_tournament(ddla, oracle, prec, n, nrhs):
    kappa = 1e4
    A = conditioned(n, kappa, 31)
    info, X, B = run(ddla, ddla.gesv_mpir, A, nrhs, prec, pivoting=ddla.PIVOT_TOURNAMENT)
    assert info.status == 0, (info.status, info.iters, info.backward_error, info.fallback)
    assert info.fallback == 0, (info.iters, info.backward_error)
    assert info.backward_error <= 1e-14
    Ao = A.copy()
    _, piv = oracle.getrf(Ao)
    Xo = oracle.getrs(Ao, piv, B.copy())
    assert rel_fro(X, Xo) <= kappa * np.sqrt(n) * 1e-14


def test_gesv_mpir_tournament_singular(ddla):
    n = 600
    A = np.random.default_rng(3).uniform(-1, 1, (n, n))
    A[:, 100] = 0.0
    info, _, _ = run(ddla, ddla.gesv_mpir, A, 1, 2, pivoting=ddla.PIVOT_TOURNAMENT)
    assert info.status == 101 and info.fallback == 3
