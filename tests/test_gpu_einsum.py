"""GPU parity of the binary tensor contraction (MArray.Contract) against the oracle's
explicit-loop restatement (small cases) and numpy.einsum, which the oracle matches on every
form einsum can state (larger cases): within 1e-12 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ("ij,jk->ik", (4, 5), (5, 3)),
    ("ij,jk->ki", (4, 5), (5, 3)),
    ("bij,bjk->bik", (3, 4, 5), (3, 5, 2)),
    ("abcd,cdef->abef", (2, 3, 4, 2), (4, 2, 3, 2)),
    ("iij,jk->ik", (4, 4, 3), (3, 2)),
    ("ijx,jk->ik", (3, 4, 2), (4, 5)),
    ("ij,kl->ijkl", (2, 3), (2, 2)),
    ("i,i->", (7,), (7,)),
    ("ij,jk->ikz", (6, 5), (5, 3)),
    ("ij,ji->ii", (4, 3), (3, 4)),
]


def _result_shape(expr, A, B):
    lhs, res = expr.split("->")
    sa, sb = lhs.split(",")
    lens = dict(zip(sb, B.shape))
    lens.update(zip(sa, A.shape))
    lens.setdefault("z", 2)
    return tuple(lens[ch] for ch in res)


@pytest.mark.parametrize("expr,sa,sb", CASES)
def test_contract_matches_oracle(ddla, oracle, expr, sa, sb):
    rng = np.random.default_rng(len(expr) + len(sa))
    A = rng.uniform(-1, 1, sa)
    B = rng.uniform(-1, 1, sb)
    C0 = rng.uniform(-1, 1, _result_shape(expr, A, B))
    Co = C0.copy()
    oracle.contract(expr, 0.75, A, B, 0.5, Co)
    C = C0.copy()
    ddla.Contract(C, expr, 0.75, A, B, beta=0.5)
    np.testing.assert_allclose(C, Co, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("expr,sa,sb", [
    ("abcd,cdef->abef", (16, 24, 32, 8), (32, 8, 20, 12)),
    ("bij,bjk->bik", (8, 300, 200), (8, 200, 250)),
    ("ikj,jl->lik", (64, 48, 96), (96, 80)),
    ("ij,jk->ik", (1024, 1024), (1024, 1024)),
])
def test_contract_larger_matches_einsum(ddla, expr, sa, sb):
    rng = np.random.default_rng(sum(sa))
    A = rng.uniform(-1, 1, sa)
    B = rng.uniform(-1, 1, sb)
    ref = np.einsum(expr, A, B)
    C = ddla.ContractNew(expr, 1.0, A, B)
    assert C.shape == ref.shape
    assert np.max(np.abs(C - ref)) <= 1e-12 * np.max(np.abs(ref)) * max(1, A.shape[-1])


def test_contract_strided_views_and_errors(ddla, oracle):
    rng = np.random.default_rng(9)
    big = rng.uniform(-1, 1, (6, 10))
    A = big[:, ::2]
    B = rng.uniform(-1, 1, (3, 5)).T
    Cbig = np.zeros((6, 6))
    C = Cbig[:, ::2]  # the result is a view with column stride 2
    ddla.Contract(C, "ij,jk->ik", 1.0, A, B, beta=0.0)
    np.testing.assert_allclose(Cbig[:, ::2], A @ B, rtol=1e-13)
    np.testing.assert_array_equal(Cbig[:, 1::2], 0.0)
    # alpha = 0: C only scaled
    D = np.ones((6, 3))
    ddla.Contract(D, "ij,jk->ik", 0.0, A, B, beta=3.0)
    np.testing.assert_array_equal(D, 3.0)
    X = rng.uniform(-1, 1, (4, 3))
    for bad in ("ij,jk", "ij->ik", "ij,jk,kl->il", "i,jk->ik"):
        with pytest.raises(ddla.DDLAError):
            ddla.Contract(np.zeros((4, 4)), bad, 1.0, X, rng.uniform(-1, 1, (3, 4)))
    with pytest.raises(ddla.DDLAError, match="different lengths"):
        ddla.Contract(np.zeros((4, 4)), "ij,jk->ik", 1.0, X, rng.uniform(-1, 1, (4, 4)))
