"""Batched uneven factorizations (one persistent graph-captured launch) against the CPU
oracle, problem by problem: LU with partial pivoting and lower Cholesky, FP64 and FP32."""
import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu

EDGE = [1, 2, 3, 8, 15, 16, 17, 31, 32, 33, 63, 64, 65, 127, 128, 129, 200, 255, 256]


def make_batch(oracle, sizes, kinds, seed):
    mats = []
    for i, (n, k) in enumerate(zip(sizes, kinds)):
        A = oracle.uniform_matrix(n, n, seed + 7 * i)
        if k == 1:  # SPD: G G^T + n I
            A = A @ A.T + n * np.eye(n)
        mats.append(A)
    return mats


def check_fp64(ddla, oracle, mats, kinds, F, P, info):
    for i, (A, k) in enumerate(zip(mats, kinds)):
        n = A.shape[0]
        assert info[i] == 0, (i, n, k, info[i])
        if k == ddla.BATCH_LU:
            Ao = A.copy()
            st, po = oracle.getrf(Ao)
            assert st == 0
            np.testing.assert_array_equal(P[i], po, err_msg=f"pivots of problem {i} (n={n})")
            if n < 128:  # the reference's unblocked LU: same operations, same order
                np.testing.assert_array_equal(F[i], Ao, err_msg=f"LU of problem {i} (n={n})")
            else:
                assert rel_fro(F[i], Ao) <= 1e-13 * n
        else:
            So = A.copy()
            assert oracle.potrf(So) == 0
            L = np.tril(F[i])
            if n <= 128:
                np.testing.assert_array_equal(L, np.tril(So), err_msg=f"Cholesky {i} (n={n})")
            else:
                assert rel_fro(L, np.tril(So)) <= 1e-13 * n
            np.testing.assert_array_equal(np.triu(F[i], 1), np.triu(A, 1))


def test_batched_fp64_edge_sizes(ddla, oracle):
    sizes = EDGE + EDGE
    kinds = [ddla.BATCH_LU] * len(EDGE) + [ddla.BATCH_CHOL] * len(EDGE)
    mats = make_batch(oracle, sizes, kinds, 100)
    F, P, info = ddla.batch_factor(mats, kinds, ddla.PREC_FP64)
    check_fp64(ddla, oracle, mats, kinds, F, P, info)


def test_batched_fp64_random_uneven(ddla, oracle):
    rng = np.random.default_rng(6)
    count = 300
    sizes = rng.integers(8, 257, count)
    kinds = rng.integers(0, 2, count)
    mats = make_batch(oracle, sizes, kinds, 500)
    F, P, info = ddla.batch_factor(mats, kinds, ddla.PREC_FP64, tune=True)
    check_fp64(ddla, oracle, mats, kinds, F, P, info)


def test_batched_fp32(ddla, oracle):
    rng = np.random.default_rng(8)
    count = 120
    sizes = np.concatenate([rng.integers(8, 257, count - len(EDGE)), EDGE])
    kinds = rng.integers(0, 2, sizes.size)
    mats = make_batch(oracle, sizes, kinds, 900)
    F, P, info = ddla.batch_factor([m.astype(np.float32) for m in mats], kinds, ddla.PREC_FP32)
    for i, (A, k) in enumerate(zip(mats, kinds)):
        n = A.shape[0]
        assert info[i] == 0
        A32 = A.astype(np.float32).astype(np.float64)
        Fi = F[i].astype(np.float64)
        if k == ddla.BATCH_LU:
            L = np.tril(Fi, -1) + np.eye(n)
            U = np.triu(Fi)
            PA = oracle.perm_matrix(P[i]) @ A32
            # FP32 GEPP: ||PA - LU|| / ||A|| of order n u_32 with small growth
            assert rel_fro(PA, L @ U) <= 2e-7 * n, (i, n)
            Ao = A.copy()
            _, po = oracle.getrf(Ao)
            assert np.mean(P[i] == po) >= 0.9  # rounding may flip a near-tie, rarely
        else:
            L = np.tril(Fi)
            assert rel_fro(A32, L @ L.T) <= 2e-7 * n, (i, n)
            np.testing.assert_array_equal(np.triu(F[i], 1), np.triu(A.astype(np.float32), 1))


def test_batched_info_and_replay(ddla, oracle):
    n = 40
    A = oracle.uniform_matrix(n, n, 3)
    A[:, 10] = 0.0  # zero pivot at column 10 (1-based 11)
    S = np.eye(n)
    S[25, 25] = -1.0  # Cholesky stops at minor 26
    good = oracle.uniform_matrix(n, n, 4)
    mats, kinds = [A, S, good], [ddla.BATCH_LU, ddla.BATCH_CHOL, ddla.BATCH_LU]
    F, P, info = ddla.batch_factor(mats, kinds, ddla.PREC_FP64)
    assert list(info) == [11, 26, 0]
    Go = good.copy()
    _, po = oracle.getrf(Go)
    np.testing.assert_array_equal(F[2], Go)
    np.testing.assert_array_equal(P[2], po)


def test_batched_plan_graph_replays(ddla, oracle):
    # the same plan executed several times on freshly uploaded inputs (graph replays with
    # host work and synchronisation in between) gives the eager result every time
    rng = np.random.default_rng(11)
    sizes = rng.integers(1, 257, 64)
    kinds = rng.integers(0, 2, 64)
    mats = make_batch(oracle, sizes, kinds, 1300)
    flat = np.concatenate([m.ravel() for m in mats])
    a = ddla._DevBuf(flat.nbytes)
    piv = ddla._DevBuf(int(sizes.sum()) * 4)
    info = ddla._DevBuf(64 * 4)
    plan = ddla.BatchPlan(sizes, kinds, ddla.PREC_FP64, a.ptr, piv.ptr, info.ptr)
    ref = None
    for rep in range(3):
        a.upload(flat)
        plan.execute()
        ddla.synchronize()
        out = (a.download(np.float64, flat.size), piv.download(np.int32, int(sizes.sum())),
               info.download(np.int32, 64))
        if ref is None:
            ref = out
            F, P, _ = ddla.batch_factor(mats, kinds, ddla.PREC_FP64)
            np.testing.assert_array_equal(out[0], np.concatenate([f.ravel() for f in F]))
        else:
            for x, y in zip(out, ref):
                np.testing.assert_array_equal(x, y)
    d = plan.describe()
    assert d["count"] == 64 and d["nb"] in (16, 32) and d["workgroups"] >= 1


def test_batched_tuning_cache(ddla, oracle, tmp_path, monkeypatch):
    # tune = 1 times the candidates once per bucket and stores the winner; the next plan of the
    # same bucket takes it from the cache without timing; tune = 2 times again
    path = tmp_path / "tune.json"
    monkeypatch.setenv("DDLA_TUNE_CACHE", str(path))
    rng = np.random.default_rng(12)
    count = 200
    sizes = rng.integers(8, 129, count)
    kinds = rng.integers(0, 2, count)
    mats = make_batch(oracle, sizes, kinds, 1200)
    n = np.array(sizes, dtype=np.int32)
    flat = np.concatenate([m.ravel() for m in mats])
    a = ddla._DevBuf(flat.nbytes)
    piv = ddla._DevBuf(int(n.sum()) * 4)
    info = ddla._DevBuf(count * 4)
    a.upload(flat)
    first = ddla.BatchPlan(n, kinds, ddla.PREC_FP64, a.ptr, piv.ptr, info.ptr, tune=1).describe()
    assert first["from_cache"] is False and len(first["tune_ms"]) >= 2
    text = path.read_text()
    assert first["tune_key"] in text and "|f64|n128|c7|mixed" in first["tune_key"]
    second = ddla.BatchPlan(n, kinds, ddla.PREC_FP64, a.ptr, piv.ptr, info.ptr, tune=1)
    d2 = second.describe()
    assert d2["from_cache"] is True and d2["tune_ms"] == []
    assert (d2["nb"], d2["wg_per_cu"]) == (first["nb"], first["wg_per_cu"])
    second.execute()
    ddla.synchronize()
    out = a.download(np.float64, flat.size)
    pv = piv.download(np.int32, int(n.sum()))
    st = info.download(np.int32, count)
    F, P, o, po = [], [], 0, 0
    for i, ni in enumerate(n):
        F.append(out[o:o + ni * ni].reshape(ni, ni).copy())
        P.append(pv[po:po + ni].copy() if kinds[i] == ddla.BATCH_LU else None)
        o += ni * ni
        po += ni
    check_fp64(ddla, oracle, mats, kinds, F, P, st)
    third = ddla.BatchPlan(n, kinds, ddla.PREC_FP64, a.ptr, piv.ptr, info.ptr, tune=2).describe()
    assert third["from_cache"] is False and len(third["tune_ms"]) >= 2
    # an unwritable cache location only disables the cache
    monkeypatch.setenv("DDLA_TUNE_CACHE", "/proc/ddla_no_such_dir/tune.json")
    d4 = ddla.BatchPlan(n, kinds, ddla.PREC_FP64, a.ptr, piv.ptr, info.ptr, tune=1).describe()
    assert d4["from_cache"] is False


@pytest.mark.parametrize("tune", [False, True])
def test_batched_fp32_lu_matches_fp32_oracle_bitwise(ddla, oracle, tune):
    # The batched LU takes every element through the unblocked elimination's operations in the
    # same order (substitution and update as fma chains in k order), so in FP32 it reproduces
    # PLUDecUnblock in FP32 (oracle.sgetf2 on the whole n x n matrix) bit for bit, pivots
    # included, at every order and panel width.
    rng = np.random.default_rng(31 + int(tune))
    sizes = np.array([1, 2, 8, 15, 16, 17, 31, 32, 33, 64, 100, 127, 128, 129, 200, 255, 256])
    mats = [rng.uniform(-1, 1, (n, n)).astype(np.float32) for n in sizes]
    kinds = [ddla.BATCH_LU] * len(mats)
    F, P, info = ddla.batch_factor(mats, kinds, ddla.PREC_FP32, tune=tune)
    for i, A in enumerate(mats):
        Q = A.copy()
        info_o, piv_o = oracle.sgetf2(Q)
        assert info[i] == info_o == 0
        np.testing.assert_array_equal(P[i], piv_o, err_msg=f"n={A.shape[0]}")
        np.testing.assert_array_equal(F[i], Q, err_msg=f"n={A.shape[0]}")
