"""The mixed LU's FP32 panel (partial pivoting) on the GPU against the oracle's FP32
restatement of PLUDecUnblock (LU.cs:344-374): pivots and factored panel bit for bit, for one
workgroup up to the largest panel (64 workgroups), ragged row counts and narrow panels, and
inside the whole FP32 factor (first panel, rows followed through the later interchanges)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _panel(m, nb, seed):
    return np.random.default_rng(seed).uniform(-1, 1, (m, nb)).astype(np.float32)


# 128..512 rows: one workgroup; 513..: several; 16384: 32; 20000 / 32768: 40 / 64 (8 slots
# per wave)
@pytest.mark.parametrize("m", [128, 129, 300, 512, 513, 1000, 1537, 4100, 8320, 16384, 20000,
                               32768])
def test_fp32_panel_matches_oracle_bitwise(ddla, oracle, m):
    P = _panel(m, 128, m)
    LU, piv, info = ddla.sgetf2(P)
    Q = P.copy()
    info_o, piv_o = oracle.sgetf2(Q)
    assert info == info_o == 0
    np.testing.assert_array_equal(piv, piv_o)
    np.testing.assert_array_equal(LU, Q)


@pytest.mark.parametrize("m,nb", [(1, 1), (7, 5), (64, 64), (200, 17), (700, 100), (5000, 33),
                                  (9000, 127)])
def test_fp32_panel_narrow_and_ragged(ddla, oracle, m, nb):
    P = _panel(m, nb, 7 * m + nb)
    LU, piv, info = ddla.sgetf2(P)
    Q = P.copy()
    info_o, piv_o = oracle.sgetf2(Q)
    assert info == info_o == 0
    np.testing.assert_array_equal(piv, piv_o)
    np.testing.assert_array_equal(LU, Q)


@pytest.mark.parametrize("m", [300, 3000])
def test_fp32_panel_ties_go_to_the_first_row(ddla, oracle, m):
    # small integers: many equal |x| in every column, exact arithmetic for a while
    P = np.random.default_rng(m).integers(-3, 4, (m, 128)).astype(np.float32)
    LU, piv, info = ddla.sgetf2(P)
    Q = P.copy()
    info_o, piv_o = oracle.sgetf2(Q)
    assert info == info_o
    np.testing.assert_array_equal(piv, piv_o)
    np.testing.assert_array_equal(LU, Q)


@pytest.mark.parametrize("m", [256, 2000])
def test_fp32_panel_zero_pivot_stops_where_the_oracle_does(ddla, oracle, m):
    # an exactly zero column stays zero under the updates: the panel stops at column 40
    P = _panel(m, 128, m + 1)
    P[:, 40] = 0.0
    Q = P.copy()
    info_o, piv_o = oracle.sgetf2(Q)
    LU, piv, info = ddla.sgetf2(P)
    assert info == info_o == 41
    np.testing.assert_array_equal(piv[:41], piv_o[:41])
    np.testing.assert_array_equal(LU, Q)


def test_fp32_panel_nan_entries(ddla, oracle):
    # a NaN counts below every number in the arg-max (AMax); NaN rows stay NaN
    P = _panel(1500, 128, 11)
    P[[5, 700, 1400], 0] = np.nan
    P[900, 50] = np.nan
    LU, piv, info = ddla.sgetf2(P)
    Q = P.copy()
    info_o, piv_o = oracle.sgetf2(Q)
    assert info == info_o
    np.testing.assert_array_equal(piv, piv_o)
    np.testing.assert_array_equal(LU, Q)


@pytest.mark.parametrize("n", [300, 1300, 4100])
def test_first_panel_of_the_fp32_factor(ddla, oracle, n):
    A = np.random.default_rng(n + 5).uniform(-1, 1, (n, n)).astype(np.float32)
    LU, perm, info = ddla.sgetrf_mp(A, ddla.PREC_FP32, ddla.PIVOT_PARTIAL)
    assert info == 0
    nb = min(128, n)
    P = A[:, :nb].copy()
    info_o, piv_o = oracle.sgetf2(P)
    assert info_o == 0
    idx = np.arange(n)
    for i, d in enumerate(piv_o):
        idx[i], idx[i + d] = idx[i + d], idx[i]
    np.testing.assert_array_equal(perm[:nb], idx[:nb])
    np.testing.assert_array_equal(LU[:nb, :nb], P[:nb])
    # rows below the panel, followed through the later panels' interchanges
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    np.testing.assert_array_equal(LU[inv[idx[nb:]], :nb], P[nb:])


# The FP64 panel (PLUDecUnblock, LU.cs:344-374) through the FP64 LU: the pivot sequence of the
# first panel and its factored columns bit for bit against the oracle's restatement, for one
# workgroup up to 32 (m <= 8192) and beyond (the column-per-lane kernel).
@pytest.mark.parametrize("n", [128, 129, 255, 256, 257, 1000, 2049, 4100, 8192, 8320])
def test_fp64_first_panel_matches_oracle_bitwise(ddla, oracle, n):
    A = np.random.default_rng(n + 17).uniform(-1, 1, (n, min(n, 128)))
    P = A.copy()
    info_o, piv_o = oracle.getf2(P)
    assert info_o == 0
    dP = ddla.DeviceMatrix.from_host(np.ascontiguousarray(A))
    piv = ddla.PLUDecUnblock(dP)
    np.testing.assert_array_equal(piv, piv_o)
    np.testing.assert_array_equal(dP.to_host(), P)
