"""The no-allocation FP64 LU (ddla_dgetrf_noalloc_dev) against ddla_dgetrf_dev through the C
ABI: same pivots, bit-identical factors, on two calls in a row (the second at a smaller order,
so the shared static scratch carries the first call's words), and a clear error above the
static scratch's order limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def factor(ddla, fn, A0):
    n = A0.shape[0]
    d = ddla.DeviceMatrix.from_host(A0)
    dpiv = ddla._DevInts(n)
    info = fn(n, d.ptr, d.ld, dpiv.ptr)
    return info, d.to_host(), dpiv.to_host()


def test_noalloc_lu_matches_allocating_entry_point(ddla, oracle):
    lib = ddla.native.lib()
    for n in (2500, 700):
        A0 = oracle.uniform_matrix(n, n, 0xDD1A0002 + n)
        i0, LU0, p0 = factor(ddla, lib.ddla_dgetrf_dev, A0)
        i1, LU1, p1 = factor(ddla, lib.ddla_dgetrf_noalloc_dev, A0)
        assert i0 == i1 == 0
        np.testing.assert_array_equal(p1, p0)
        np.testing.assert_array_equal(LU1, LU0)
    # a zero pivot reports the same info value
    A = oracle.uniform_matrix(300, 300, 5)
    A[:, 200] = 0.0
    assert factor(ddla, lib.ddla_dgetrf_noalloc_dev, A)[0] == 201


def test_noalloc_lu_rejects_orders_above_its_scratch(ddla):
    lib = ddla.native.lib()
    st = lib.ddla_dgetrf_noalloc_dev(32769, None, 32769, None)
    assert st == ddla.native.ERR_INVALID_ARG
    assert "32768" in ddla.native.last_error()
