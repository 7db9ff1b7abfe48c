"""DDLA's dense hot path on MI355X: Python mirror of DDLA's GEMM / LU / Cholesky API over
the C ABI of libddla_hip.so (hand-written HIP kernels for gfx950).

Names, argument meaning and error behaviour follow the reference C# API:
  BlasProvider.GeMM / TrSM / SyRk / AMax   DDLA/BLAS/3m.cs:10-314, DDLA/BLAS/1v.cs:30-41
  LU, PLUDecompose, PLUSolve, ApplyPiv     DDLA/Factorizations/LU.cs:13-465
  Cholesky, CholDecompose, CholeskySolve   DDLA/Factorizations/Cholesky.cs:10-236
Matrices are numpy arrays on the host (any strides, like MatrixView) or DeviceMatrix
objects that stay resident in HBM.  There is no CPU fallback: every call runs on the GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import native
from .native import (LEFT, LOWER, NO_TRANSPOSE, NONUNIT_DIAG, RIGHT, TRANSPOSE, UNIT_DIAG, UPPER,
                     PREC_BF16, PREC_FP16, PREC_FP32, PREC_FP64, DDLAError, MpirInfo, MpirOpts,
                     check)

__all__ = [
    "DeviceMatrix", "BlasProvider", "LU", "Cholesky", "PLUDecompose", "PLUSolve", "ApplyPiv",
    "CholDecompose", "CholeskySolve", "LinalgException", "gesv_mpir", "posv_mpir", "init",
    "synchronize", "NO_TRANSPOSE", "TRANSPOSE", "LOWER", "UPPER", "UNIT_DIAG", "NONUNIT_DIAG",
    "LEFT", "RIGHT", "PREC_FP64", "PREC_FP32", "PREC_FP16", "PREC_BF16", "DDLAError",
    "BatchPlan", "batch_factor", "BATCH_LU", "BATCH_CHOL", "sgetrf_mp", "PIVOT_PARTIAL",
    "PIVOT_BUTTERFLY", "PIVOT_TOURNAMENT", "LDLT", "LDLTDecompose", "LDLTSolve", "QR",
    "QRDecompose", "QRApply", "QRSolve", "qr_t_rows", "Contract", "ContractNew", "contract_dev",
]


class LinalgException(ArithmeticError):
    """Non-positive Cholesky pivot (DDLA/Core/LinalgException.cs, Cholesky.cs:228-236)."""


def _lib():
    return native.lib()


def init(device: int = 0):
    check(_lib().ddla_init(device))


def synchronize():
    check(_lib().ddla_sync())


def finalize():
    """Releases the library context (also done at interpreter exit)."""
    native.finalize()


def stream_handle() -> int:
    """The library's hipStream_t as an integer (for torch.cuda.ExternalStream)."""
    p = ctypes.c_void_p()
    check(_lib().ddla_get_stream(ctypes.byref(p)))
    return p.value or 0


def profile_enable(on: bool = True):
    check(_lib().ddla_profile_enable(1 if on else 0))


def profile_only(name=None):
    """Restricts profiling to one kernel name (None: every kernel)."""
    check(_lib().ddla_profile_only(name.encode() if name else None))


def profile_reset():
    check(_lib().ddla_profile_reset())


def profile_read() -> dict:
    """{kernel: {count, ms, flops, bytes}} accumulated while profiling was enabled."""
    L = _lib()
    n = L.ddla_profile_read(-1, None, 0, None, None, None, None)
    out = {}
    for i in range(n):
        name = ctypes.create_string_buffer(128)
        cnt, ms, fl, by = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(L.ddla_profile_read(i, name, 128, ctypes.byref(cnt), ctypes.byref(ms),
                                  ctypes.byref(fl), ctypes.byref(by)))
        out[name.value.decode()] = {"count": cnt.value, "ms": ms.value, "flops": fl.value,
                                    "bytes": by.value}
    return out


def _as_f64(a):
    a = np.asarray(a)
    if a.dtype != np.float64:
        raise TypeError("DDLA matrices are float64")
    return a


def _strides(a):
    return a.strides[0] // 8, a.strides[1] // 8


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class DeviceMatrix:
    """Row-major FP64 matrix resident in device memory (element (i,j) at ptr[i*ld + j])."""

    def __init__(self, rows: int, cols: int, ld: int | None = None):
        self.rows, self.cols = int(rows), int(cols)
        self.ld = int(ld) if ld is not None else max(self.cols, 1)
        p = ctypes.c_void_p()
        check(_lib().ddla_malloc(ctypes.byref(p), max(self.rows * self.ld, 1) * 8))
        self._ptr = p
        self._owner = True

    @property
    def ptr(self) -> ctypes.c_void_p:
        return self._ptr

    @property
    def shape(self):
        return (self.rows, self.cols)

    @classmethod
    def from_host(cls, a) -> "DeviceMatrix":
        a = _as_f64(a)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        d = cls(a.shape[0], a.shape[1])
        d.set(a)
        return d

    @classmethod
    def wrap(cls, ptr: int, rows: int, cols: int, ld: int, keepalive=None) -> "DeviceMatrix":
        """A non-owning view of device memory allocated elsewhere (e.g. a torch tensor)."""
        v = object.__new__(cls)
        v.rows, v.cols, v.ld = int(rows), int(cols), int(ld)
        v._ptr = ctypes.c_void_p(int(ptr))
        v._owner = False
        v._base = keepalive
        return v

    @classmethod
    def random(cls, rows: int, cols: int, seed: int) -> "DeviceMatrix":
        """SplitMix64 uniform[-1,1) fill on the device (same stream as the CPU oracle)."""
        d = cls(rows, cols)
        check(_lib().ddla_drandm_dev(d.rows, d.cols, d.ptr, d.ld, ctypes.c_uint64(seed)))
        return d

    def set(self, a):
        a = _as_f64(a)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        assert a.shape == self.shape
        check(_lib().ddla_dmat_set(self.rows, self.cols, _ptr(a), *_strides(a), self.ptr, self.ld))

    def to_host(self) -> np.ndarray:
        out = np.empty((self.rows, self.cols), dtype=np.float64)
        check(_lib().ddla_dmat_get(self.rows, self.cols, self.ptr, self.ld, _ptr(out), *_strides(out)))
        return out

    def sub(self, i: int, m: int, j: int, n: int) -> "DeviceMatrix":
        """A non-owning view of rows [i, i+m) and columns [j, j+n)."""
        v = object.__new__(DeviceMatrix)
        v.rows, v.cols, v.ld = m, n, self.ld
        v._ptr = ctypes.c_void_p(self._ptr.value + (i * self.ld + j) * 8)
        v._owner = False
        v._base = self
        return v

    def free(self):
        if getattr(self, "_owner", False) and self._ptr:
            if not native.finalized():
                _lib().ddla_free(self._ptr)
            self._ptr = ctypes.c_void_p()
            self._owner = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BlasProvider:
    """GPU replacements for the BLAS entry points on DDLA's dense path."""

    @staticmethod
    def GeMM(alpha, A, B, beta, C, transA=NO_TRANSPOSE, transB=NO_TRANSPOSE):
        """C = alpha*op(A)*op(B) + beta*C (DDLA/BLAS/3m.cs:10-42 -> bli_dgemm)."""
        al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
        if isinstance(C, DeviceMatrix):
            m, n = C.shape
            k = A.rows if transA == TRANSPOSE else A.cols
            return check(_lib().ddla_dgemm_dev(transA, transB, m, n, k, ctypes.byref(al), A.ptr,
                                               A.ld, 1, B.ptr, B.ld, 1, ctypes.byref(be), C.ptr,
                                               C.ld, 1))
        A, B, C = _as_f64(A), _as_f64(B), _as_f64(C)
        m, n = C.shape
        k = A.shape[0] if transA == TRANSPOSE else A.shape[1]
        ka, kb = (A.shape[1] if transA == TRANSPOSE else A.shape[0]), \
            (B.shape[0] if transB == TRANSPOSE else B.shape[1])
        if ka != m or kb != n or (B.shape[1] if transB == TRANSPOSE else B.shape[0]) != k:
            raise ValueError("Dimensions of matrices A, B and C must match")
        return check(_lib().ddla_dgemm(transA, transB, m, n, k, ctypes.byref(al), _ptr(A),
                                       *_strides(A), _ptr(B), *_strides(B), ctypes.byref(be),
                                       _ptr(C), *_strides(C)))

    @staticmethod
    def TrSM(side, uplo, trans, diag, alpha, A, B):
        """B = alpha*op(A)^-1*B or alpha*B*op(A)^-1 (DDLA/BLAS/3m.cs:287-305).
        numpy operands (any strides) go through bli_dtrsm's argument list (ddla_dtrsm)."""
        al = ctypes.c_double(alpha)
        m, n = _shape(B)
        _check_square(A, m if side == LEFT else n)
        if isinstance(B, DeviceMatrix):
            return check(_lib().ddla_dtrsm_dev(side, uplo, trans, diag, m, n, ctypes.byref(al),
                                               A.ptr, A.ld, B.ptr, B.ld))
        return check(_lib().ddla_dtrsm(side, uplo, trans, diag, m, n, ctypes.byref(al),
                                       *_view(A), *_view(B)))

    @staticmethod
    def SyRk(uplo, trans, alpha, A, beta, C):
        """uplo(C) = beta*C + alpha*op(A)*op(A)^T (DDLA/BLAS/3m.cs:118-138)."""
        al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
        m = _check_square(C)
        ma, k = _op_shape(A, trans)
        if ma != m:
            raise ValueError("Dimensions of matrixs A must be match!")
        if isinstance(C, DeviceMatrix):
            return check(_lib().ddla_dsyrk_dev(uplo, trans, m, k, ctypes.byref(al), A.ptr, A.ld,
                                               ctypes.byref(be), C.ptr, C.ld))
        return check(_lib().ddla_dsyrk(uplo, trans, m, k, ctypes.byref(al), *_view(A),
                                       ctypes.byref(be), *_view(C)))

    @staticmethod
    def GeMMT(uplo, transA, transB, alpha, A, B, beta, C):
        """uplo(C) = beta*C + alpha*op(A)*op(B) (DDLA/BLAS/3m.cs:44-70 -> bli_dgemmt)."""
        al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
        m = _check_square(C)
        ma, k = _op_shape(A, transA)
        kb, mb = _op_shape(B, transB)
        if ma != m or kb != k or mb != m:
            raise ValueError("Dimensions of matrix A must be match!")
        fn = _lib().ddla_dgemmt_dev if isinstance(C, DeviceMatrix) else _lib().ddla_dgemmt
        return check(fn(uplo, transA, transB, m, k, ctypes.byref(al), *_view(A), *_view(B),
                        ctypes.byref(be), *_view(C)))

    @staticmethod
    def SyR2k(uplo, transA, transB, alpha, A, B, beta, C):
        """uplo(C) = beta*C + alpha*(op(A) op(B)^T + op(B) op(A)^T) (DDLA/BLAS/3m.cs:149-172)."""
        al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
        m = _check_square(C)
        ma, k = _op_shape(A, transA)
        mb, kb = _op_shape(B, transB)
        if ma != m or mb != m or kb != k:
            raise ValueError("Dimensions of matrixs A and B must be match!")
        fn = _lib().ddla_dsyr2k_dev if isinstance(C, DeviceMatrix) else _lib().ddla_dsyr2k
        return check(fn(uplo, transA, transB, m, k, ctypes.byref(al), *_view(A), *_view(B),
                        ctypes.byref(be), *_view(C)))

    @staticmethod
    def SyMM(side, uplo, transB, alpha, A, B, beta, C):
        """C = beta*C + alpha*A*op(B) (LEFT) or alpha*op(B)*A (RIGHT), A symmetric from its
        uplo triangle (DDLA/BLAS/3m.cs:83-106 -> bli_dsymm)."""
        al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
        m, n = _shape(C)
        if _op_shape(B, transB) != (m, n):
            raise ValueError("Dimensions of matrixs must be match!")
        _check_square(A, m if side == LEFT else n)
        fn = _lib().ddla_dsymm_dev if isinstance(C, DeviceMatrix) else _lib().ddla_dsymm
        return check(fn(side, uplo, 0, transB, m, n, ctypes.byref(al), *_view(A), *_view(B),
                        ctypes.byref(be), *_view(C)))

    @staticmethod
    def TrMM(side, uplo, trans, diag, alpha, A, B):
        """B = alpha*op(A)*B (LEFT) or alpha*B*op(A) (RIGHT), A triangular
        (DDLA/BLAS/3m.cs:185-204 -> bli_dtrmm)."""
        al = ctypes.c_double(alpha)
        m, n = _shape(B)
        _check_square(A, m if side == LEFT else n)
        fn = _lib().ddla_dtrmm_dev if isinstance(B, DeviceMatrix) else _lib().ddla_dtrmm
        return check(fn(side, uplo, trans, diag, m, n, ctypes.byref(al), *_view(A), *_view(B)))

    @staticmethod
    def TrMM3(side, uplo, transA, diag, transB, alpha, A, B, beta, C):
        """C = beta*C + alpha*op(A)*op(B) (LEFT) or alpha*op(B)*op(A) (RIGHT), A triangular
        (DDLA/BLAS/3m.cs:225-250 -> bli_dtrmm3)."""
        al, be = ctypes.c_double(alpha), ctypes.c_double(beta)
        m, n = _shape(C)
        if _op_shape(B, transB) != (m, n):
            raise ValueError("Dimensions of matrixs must be match!")
        _check_square(A, m if side == LEFT else n)
        fn = _lib().ddla_dtrmm3_dev if isinstance(C, DeviceMatrix) else _lib().ddla_dtrmm3
        return check(fn(side, uplo, transA, diag, transB, m, n, ctypes.byref(al), *_view(A),
                        *_view(B), ctypes.byref(be), *_view(C)))

    @staticmethod
    def AMax(x: DeviceMatrix, inc: int = 1) -> int:
        """Index of the first max |x_i| of a device vector (DDLA/BLAS/1v.cs:30-41)."""
        n = x.rows * x.cols if inc == 1 else x.rows
        out = ctypes.c_int64()
        check(_lib().ddla_damaxv_dev(n, x.ptr, inc if inc != 1 else 1, ctypes.byref(out)))
        return int(out.value)


def _shape(X):
    return X.shape if isinstance(X, DeviceMatrix) else _as_f64(X).shape


def _op_shape(X, trans):
    r, c = _shape(X)
    return (c, r) if trans == TRANSPOSE else (r, c)


def _check_square(X, n=None):
    r, c = _shape(X)
    if r != c or (n is not None and r != n):
        raise ValueError("Dimensions of matrixs A must be match!")
    return r


def _view(X):
    """(pointer, row stride, column stride) of a DeviceMatrix or a numpy array (elements)."""
    if isinstance(X, DeviceMatrix):
        return X.ptr, X.ld, 1
    X = _as_f64(X)
    return (_ptr(X), *_strides(X))


def _zero_pivot(info: int):
    # reference: InvScal throws ArgumentOutOfRangeException for a zero pivot (1v.cs:99)
    raise ZeroDivisionError(f"zero pivot at step {info - 1}")


def PLUDecUnblock(A: "DeviceMatrix"):
    """PLUDecUnblock (LU.cs:344-374) on a device panel in place: A is m x nb with
    1 <= nb <= min(m, 128).  Returns the relative pivots (int32[nb]); raises on a zero pivot
    like the reference (the panel stops there)."""
    m, nb = A.rows, A.cols
    dpiv = _DevInts(nb)
    info = check(_lib().ddla_dgetf2_dev(m, nb, A.ptr, A.ld, dpiv.ptr))
    piv = dpiv.to_host()
    if info:
        _zero_pivot(info)
    return piv


def PLUDecompose(A, pivs=None):
    """Blocked PLU in place (LU.cs:251-264). A: numpy (any strides) or DeviceMatrix.

    Returns the relative pivots (int32[n]); raises on a zero pivot like the reference."""
    if isinstance(A, DeviceMatrix):
        n = A.rows
        if A.cols != n:
            raise ValueError("Matrix must be square for LU factorization.")
        dpiv = _DevInts(n)
        info = check(_lib().ddla_dgetrf_dev(n, A.ptr, A.ld, dpiv.ptr))
        piv = dpiv.to_host()
        if pivs is not None:
            pivs[:] = piv
        if info:
            _zero_pivot(info)
        return piv
    A = _as_f64(A)
    n = A.shape[0]
    if A.shape[1] != n:
        raise ValueError("Matrix must be square for LU factorization.")
    piv = np.zeros(n, dtype=np.int32) if pivs is None else pivs
    if piv.shape != (n,) or piv.dtype != np.int32:
        raise ValueError("The length of pivs must match the number of rows in the matrix.")
    info = check(_lib().ddla_dgetrf(n, _ptr(A), *_strides(A), _ptr(piv)))
    if info:
        _zero_pivot(info)
    return piv


def PLUSolve(A, pivs, B, trans=False):
    """B <- A^-1 B (or A^-T B) from the packed factors (LU.cs:266-287)."""
    if isinstance(A, DeviceMatrix):
        dpiv = pivs if isinstance(pivs, _DevInts) else _DevInts.from_host(pivs)
        return check(_lib().ddla_dgetrs_dev(TRANSPOSE if trans else NO_TRANSPOSE, A.rows, B.cols,
                                            A.ptr, A.ld, dpiv.ptr, B.ptr, B.ld))
    A, B = _as_f64(A), _as_f64(B)
    if B.ndim == 1:
        B2 = B.reshape(-1, 1)
    else:
        B2 = B
    if A.shape[0] != B2.shape[0]:
        raise ValueError("RHS matrix size mismatch")
    piv = np.ascontiguousarray(pivs, dtype=np.int32)
    return check(_lib().ddla_dgetrs(TRANSPOSE if trans else NO_TRANSPOSE, A.shape[0], B2.shape[1],
                                    _ptr(A), *_strides(A), _ptr(piv), _ptr(B2), *_strides(B2)))


def LUDecompose(A):
    """LU without pivoting in place (LU.cs:133-142). A: numpy or DeviceMatrix."""
    d = A if isinstance(A, DeviceMatrix) else DeviceMatrix.from_host(_as_f64(A))
    if d.rows != d.cols:
        raise ValueError("Matrix must be square for LU factorization.")
    info = check(_lib().ddla_dgetrf_nopiv_dev(d.rows, d.ptr, d.ld))
    if not isinstance(A, DeviceMatrix):
        A[...] = d.to_host()
    if info:
        _zero_pivot(info)
    return 0


def LUSolve(A, B, trans=False):
    """B <- A^-1 B (or A^-T B) from LU factors without pivoting (LU.cs:144-152)."""
    dA = A if isinstance(A, DeviceMatrix) else DeviceMatrix.from_host(_as_f64(A))
    dB = B if isinstance(B, DeviceMatrix) else DeviceMatrix.from_host(_as_f64(B))
    if dB.rows != dA.rows:
        raise ValueError("RHS matrix size mismatch")
    check(_lib().ddla_dgetrs_nopiv_dev(TRANSPOSE if trans else NO_TRANSPOSE, dA.rows, dB.cols,
                                       dA.ptr, dA.ld, dB.ptr, dB.ld))
    if not isinstance(B, DeviceMatrix):
        B[...] = dB.to_host().reshape(B.shape)
    return 0


def ApplyPiv(B: DeviceMatrix, pivs, side=LEFT, forward=True):
    """Row (left) or column (right) interchanges with relative pivots (LU.cs:237-249)."""
    npiv = len(pivs) if not isinstance(pivs, _DevInts) else pivs.n
    dpiv = pivs if isinstance(pivs, _DevInts) else _DevInts.from_host(pivs)
    return check(_lib().ddla_dlaswp_dev(side, 1 if forward else 0, B.rows, B.cols, B.ptr, B.ld,
                                        dpiv.ptr, npiv))


class _DevInts:
    def __init__(self, n):
        self.n = int(n)
        p = ctypes.c_void_p()
        check(_lib().ddla_malloc(ctypes.byref(p), max(self.n, 1) * 4))
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        d = cls(a.size)
        check(_lib().ddla_memcpy_h2d(d.ptr, _ptr(a), a.size * 4))
        return d

    def to_host(self):
        out = np.empty(self.n, dtype=np.int32)
        check(_lib().ddla_memcpy_d2h(_ptr(out), self.ptr, self.n * 4))
        return out

    def __del__(self):
        try:
            if self.ptr and not native.finalized():
                _lib().ddla_free(self.ptr)
        except Exception:
            pass


class LU:
    """LU with partial pivoting (LU.cs:13-128); the factorization runs once, lazily."""

    def __init__(self, A, inplace=False, usePiv=True):
        if not usePiv:
            raise NotImplementedError("LU without pivoting is not on the GPU path")
        A = _as_f64(A)
        n = A.shape[0]
        if A.shape[1] != n:
            raise ValueError("A must be square")
        self.n = n
        self.matrix = A if inplace else np.array(A, order="C", copy=True)
        self.pivs = np.zeros(n, dtype=np.int32)
        self._computed = False

    def _compute_once(self):
        if not self._computed:
            PLUDecompose(self.matrix, self.pivs)
            self._computed = True

    def Solve(self, b, trans=False, inplace=False):
        """Solves A x = b (the reference's vector overload ignores trans, LU.cs:77; this
        mirror honours it)."""
        self._compute_once()
        b = _as_f64(b)
        if b.shape[0] != self.n:
            raise ValueError("RHS vector size mismatch")
        x = b if inplace else np.array(b, copy=True)
        PLUSolve(self.matrix, self.pivs, x, trans)
        return x

    def Inverse(self):
        """A^-1 from the factors (LU.Inverse, LU.cs:22-27, :102-110).  The reference builds the
        identity and calls Solve(inverse, true), which solves with the transpose into a new
        matrix and returns the untouched identity; this mirror returns the inverse."""
        self._compute_once()
        X = np.eye(self.n)
        PLUSolve(self.matrix, self.pivs, X, False)
        return X

    def Deconstruct(self):
        """(P, L, U) with P A = L U (LU.Deconstruct, LU.cs:112-128)."""
        self._compute_once()
        n = self.n
        P = np.eye(n)
        d = DeviceMatrix.from_host(P)
        ApplyPiv(d, self.pivs)
        P = d.to_host()
        L = np.tril(self.matrix, -1) + np.eye(n)
        U = np.triu(self.matrix)
        return P, L, U


def CholDecompose(uplo, A):
    """Blocked Cholesky in place (Cholesky.cs:52-69); raises LinalgException if not SPD."""
    if isinstance(A, DeviceMatrix):
        info = check(_lib().ddla_dpotrf_dev(uplo, A.rows, A.ptr, A.ld))
    else:
        A = _as_f64(A)
        if A.shape[0] != A.shape[1]:
            raise ValueError("Matrix must be square")
        info = check(_lib().ddla_dpotrf(uplo, A.shape[0], _ptr(A), *_strides(A)))
    if info:
        raise LinalgException(f"Cholesky factorization encountered a non-positive pivot at {info - 1}")
    return 0


def CholeskySolve(uplo, A, B):
    """B <- A^-1 B with a Cholesky factor (Cholesky.cs:102-123)."""
    if isinstance(A, DeviceMatrix):
        return check(_lib().ddla_dpotrs_dev(uplo, A.rows, B.cols, A.ptr, A.ld, B.ptr, B.ld))
    A, B = _as_f64(A), _as_f64(B)
    B2 = B.reshape(-1, 1) if B.ndim == 1 else B
    if B2.shape[0] != A.shape[0]:
        raise ValueError("Matrix dimensions do not match for Cholesky solve.")
    return check(_lib().ddla_dpotrs(uplo, A.shape[0], B2.shape[1], _ptr(A), *_strides(A),
                                    _ptr(B2), *_strides(B2)))


class Cholesky:
    """Cholesky factorization (Cholesky.cs:10-88)."""

    def __init__(self, A, uplo=LOWER, inplace=False):
        A = _as_f64(A)
        if A.shape[0] != A.shape[1]:
            raise ValueError("Matrix must be square")
        self.uplo = uplo
        self.matrix = A if inplace else np.array(A, order="C", copy=True)
        self._computed = False

    def ComputeOnce(self):
        if not self._computed:
            self._computed = True
            CholDecompose(self.uplo, self.matrix)

    def Solve(self, b, inplace=False):
        self.ComputeOnce()
        x = b if inplace else np.array(_as_f64(b), copy=True)
        CholeskySolve(self.uplo, self.matrix, x)
        return x


def LDLTDecompose(uplo, A):
    """Blocked LDL^T without pivoting in place (LDLT.cs:56-73): lower A = L D L^T, upper
    A = U^T D U.  Raises LinalgException on a zero pivot (the reference checks every pivot
    but the last of each 128-block, LDLT.cs:145-152; this checks all of them)."""
    if isinstance(A, DeviceMatrix):
        if A.rows != A.cols:
            raise ValueError("Matrix must be square")
        info = check(_lib().ddla_dldlt_dev(uplo, A.rows, A.ptr, A.ld))
    else:
        A = _as_f64(A)
        if A.shape[0] != A.shape[1]:
            raise ValueError("Matrix must be square")
        info = check(_lib().ddla_dldlt(uplo, A.shape[0], _ptr(A), *_strides(A)))
    if info:
        raise LinalgException(f"LDLT factorization encountered a zero pivot at {info - 1}")
    return 0


def LDLTSolve(uplo, A, B):
    """B <- A^-1 B with an LDL^T factor (LDLT.cs:105-126, :267-295)."""
    if isinstance(A, DeviceMatrix):
        if B.rows != A.rows:
            raise ValueError("Matrix dimensions do not match for Cholesky solve.")
        return check(_lib().ddla_dldlt_solve_dev(uplo, A.rows, B.cols, A.ptr, A.ld, B.ptr, B.ld))
    A, B = _as_f64(A), _as_f64(B)
    B2 = B.reshape(-1, 1) if B.ndim == 1 else B
    if B2.shape[0] != A.shape[0]:
        raise ValueError("Matrix dimensions do not match for Cholesky solve.")
    return check(_lib().ddla_dldlt_solve(uplo, A.shape[0], B2.shape[1], _ptr(A), *_strides(A),
                                         _ptr(B2), *_strides(B2)))


class LDLT:
    """LDL^T factorization (LDLT.cs:9-137); computed once, lazily."""

    def __init__(self, A, uplo=LOWER, inplace=False):
        A = _as_f64(A)
        if A.shape[0] != A.shape[1]:
            raise ValueError("Matrix must be square")
        self.uplo = uplo
        self.matrix = A if inplace else np.array(A, order="C", copy=True)
        self._computed = False
        self._deconstructed = False

    def ComputeOnce(self):
        if self._deconstructed:
            raise RuntimeError("Matrix has been deconstructed, cannot compute again.")
        if not self._computed:
            self._computed = True
            LDLTDecompose(self.uplo, self.matrix)

    def Solve(self, b, inplace=False):
        self.ComputeOnce()
        x = b if inplace else np.array(_as_f64(b), copy=True)
        LDLTSolve(self.uplo, self.matrix, x)
        return x

    def Deconstruct(self):
        """(L, D): the unit triangle of the chosen side and the diagonal matrix
        (LDLT.cs:128-137); for UPPER the triangle returned is U."""
        self.ComputeOnce()
        D = np.diag(np.diag(self.matrix).copy())
        L = np.tril(self.matrix) if self.uplo == LOWER else np.triu(self.matrix)
        np.fill_diagonal(L, 1.0)
        self._deconstructed = True
        return L, D


def qr_t_rows(m, n):
    """Rows of the T matrix of a QR (QR.CreateT, QR.cs:124-128): min(256, max(m, n))."""
    return int(_lib().ddla_qr_t_rows(m, n))


def QRDecompose(A, T=None):
    """Blocked Householder QR in the UT form, in place (QR.cs:130-149): A gets R and the
    Householder vectors; returns T (qr_t_rows(m, n) x n, the block factors).  A and T are
    numpy arrays (any strides) or DeviceMatrix objects (T required then)."""
    if isinstance(A, DeviceMatrix):
        if T is None or T.rows < qr_t_rows(A.rows, A.cols) or T.cols != A.cols:
            raise ValueError("Matrix T and A must have the same number of columns.")
        check(_lib().ddla_dgeqrf_dev(A.rows, A.cols, A.ptr, A.ld, T.ptr, T.ld))
        return T
    A = _as_f64(A)
    m, n = A.shape
    if T is None:
        T = np.zeros((qr_t_rows(m, n), n))
    if T.shape != (qr_t_rows(m, n), n):
        raise ValueError("Matrix T and A must have the same number of columns.")
    check(_lib().ddla_dgeqrf(m, n, _ptr(A), *_strides(A), _ptr(T), *_strides(T)))
    return T


def QRApply(A, T, B, trans):
    """B <- Q^T B (trans, ApplyQlhfc) or Q B (ApplyQlnfc) on DeviceMatrix operands."""
    return check(_lib().ddla_dqr_apply_dev(TRANSPOSE if trans else NO_TRANSPOSE, A.rows, A.cols,
                                           A.ptr, A.ld, T.ptr, T.ld, B.cols, B.ptr, B.ld))


def QRSolve(A, T, B, X=None):
    """Least squares X = argmin |A X - B| (QR.Solve, QR.cs:151-167); B unchanged."""
    if isinstance(A, DeviceMatrix):
        check(_lib().ddla_dqr_solve_dev(A.rows, A.cols, B.cols, A.ptr, A.ld, T.ptr, T.ld, B.ptr,
                                        B.ld, X.ptr, X.ld))
        return X
    A, T, B = _as_f64(A), _as_f64(T), _as_f64(B)
    B2 = B.reshape(-1, 1) if B.ndim == 1 else B
    m, n = A.shape
    if B2.shape[0] != m:
        raise ValueError("Matrix A and B must have the same number of rows.")
    X2 = np.zeros((n, B2.shape[1])) if X is None else X.reshape(n, -1)
    check(_lib().ddla_dqr_solve(m, n, B2.shape[1], _ptr(A), *_strides(A), _ptr(T), *_strides(T),
                                _ptr(B2), *_strides(B2), _ptr(X2), *_strides(X2)))
    return X2.reshape(-1) if B.ndim == 1 and X is None else X2


class QR:
    """Householder QR (QR.cs:14-122); computed once, lazily, on the GPU."""

    def __init__(self, A, inplace=False):
        A = _as_f64(A)
        self.matrix = A if inplace else np.array(A, order="C", copy=True)
        m, n = self.matrix.shape
        self.aux = np.zeros((qr_t_rows(m, n), n))
        self._computed = False
        self._deconstructed = False

    def ComputeOnce(self):
        if self._deconstructed:
            raise RuntimeError("Matrix has been deconstructed, cannot compute again.")
        if not self._computed:
            self._computed = True
            QRDecompose(self.matrix, self.aux)

    def Solve(self, B):
        self.ComputeOnce()
        return QRSolve(self.matrix, self.aux, B)

    def _apply(self, B, trans):
        B = _as_f64(B)
        B2 = B.reshape(-1, 1) if B.ndim == 1 else B
        m = self.matrix.shape[0]
        if B2.shape[0] != m:
            raise ValueError("B must have as many rows as A")
        dA = DeviceMatrix.from_host(self.matrix)
        dT = DeviceMatrix.from_host(self.aux)
        dB = DeviceMatrix.from_host(np.ascontiguousarray(B2))
        QRApply(dA, dT, dB, trans)
        B2[...] = dB.to_host()
        return B

    def ApplyQ(self, B):
        """B <- Q B in place (QR.cs:78-90)."""
        self.ComputeOnce()
        return self._apply(B, False)

    def ApplyQT(self, B):
        """B <- Q^T B in place (QR.cs:92-104)."""
        self.ComputeOnce()
        return self._apply(B, True)

    def Deconstruct(self):
        """(Q, R): Q = H_0 ... H_{k-1} applied to the m x m identity, R the upper triangle
        (QR.cs:106-118)."""
        self.ComputeOnce()
        m = self.matrix.shape[0]
        Q = self._apply(np.eye(m), False)
        R = np.triu(self.matrix)
        self._deconstructed = True
        return Q, R


def _dev_tensor(x):
    """A contiguous device copy of a numpy tensor: (DeviceMatrix keeping it, shape, strides)."""
    x = np.array(_as_f64(x), order="C", copy=True)  # keeps a 0-d tensor 0-d
    flat = x.reshape(1, max(1, x.size)) if x.size else np.zeros((1, 1))
    d = DeviceMatrix.from_host(flat)
    strides = [st // 8 for st in x.strides]
    return d, list(x.shape), strides


def _i64s(v):
    return (ctypes.c_int64 * max(1, len(v)))(*v)


def contract_dev(expression, alpha, a_ptr, a_shape, a_strides, b_ptr, b_shape, b_strides, beta,
                 c_ptr, c_shape, c_strides):
    """C = beta C + alpha contract(expression, A, B) on device tensors given by pointer,
    lengths and element strides (MArray.Contract, DDLA/Einsum/MArray.Contract.cs:35-125)."""
    return check(_lib().ddla_dcontract_dev(
        expression.encode(), alpha, a_ptr, len(a_shape), _i64s(a_shape), _i64s(a_strides),
        b_ptr, len(b_shape), _i64s(b_shape), _i64s(b_strides), beta, c_ptr, len(c_shape),
        _i64s(c_shape), _i64s(c_strides)))


def Contract(result, expression, alpha, left, right, beta=1.0):
    """result = beta*result + alpha*contract(expression, left, right) on the GPU
    (MArray.Contract(beta, result, expression, alpha, left, right), MArray.Contract.cs:35-58);
    numpy tensors of any strides, result updated in place."""
    C = _as_f64(result)
    dA, sa, ta = _dev_tensor(left)
    dB, sb, tb = _dev_tensor(right)
    dC, sc, tc = _dev_tensor(C)
    contract_dev(expression, alpha, dA.ptr, sa, ta, dB.ptr, sb, tb, beta, dC.ptr, sc, tc)
    if C.size:
        C[...] = dC.to_host().reshape(C.shape)
    return C


def ContractNew(expression, alpha, left, right):
    """A new result tensor (MArray.Contract(expression, alpha, left, right), :131-250): its
    lengths from the symbols' lengths in left or right."""
    lhs, res = [t.strip() for t in expression.split("->")]
    sa, sb = [t.strip() for t in lhs.split(",")]
    left, right = _as_f64(left), _as_f64(right)
    lens = {}
    for sym, arr in ((sb, right), (sa, left)):
        for ch, L in zip(sym, arr.shape):
            lens[ch] = L
    missing = [ch for ch in res if ch not in lens]
    if missing:
        raise ValueError(f"Symbol '{missing[0]}' not found in either tensor.")
    C = np.zeros([lens[ch] for ch in res])
    return Contract(C, expression, alpha, left, right, beta=0.0)


PIVOT_PARTIAL, PIVOT_BUTTERFLY, PIVOT_TOURNAMENT = 0, 1, 2


def _mpir(fn, A: DeviceMatrix, B: DeviceMatrix, X: DeviceMatrix, prec, tol, max_iters, gmres,
          terms, pivoting=PIVOT_PARTIAL, x_ref=None):
    opts = MpirOpts(prec, max_iters, tol, gmres, terms, pivoting, 0,
                    x_ref.ptr if x_ref is not None else None,
                    x_ref.ld if x_ref is not None else 0)
    info = MpirInfo()
    st = fn(A.rows, B.cols, A.ptr, A.ld, B.ptr, B.ld, X.ptr, X.ld, ctypes.byref(opts),
            ctypes.byref(info))
    if st < 0 and st != native.ERR_NOT_CONVERGED:
        check(st)
    return info


def gesv_mpir(A: DeviceMatrix, B: DeviceMatrix, X: DeviceMatrix, prec=PREC_FP16, tol=1e-14,
              max_iters=50, gmres=0, terms=0, pivoting=PIVOT_PARTIAL, x_ref=None) -> MpirInfo:
    """Mixed-precision LU + iterative refinement to FP64 accuracy (no reference counterpart).

    terms: low-precision parts per operand of the trailing updates (0 = library default).
    pivoting: PIVOT_PARTIAL; PIVOT_BUTTERFLY (random butterfly transform + LU without
    pivoting); or PIVOT_TOURNAMENT (pivot rows of each 128-column panel chosen by a reduction
    tree, as in CALU).  The last two fall back to partial pivoting if they do not refine.
    x_ref: optional reference solution (DeviceMatrix): info.forward_error is measured
    against it (else -1)."""
    return _mpir(_lib().ddla_dgesv_mpir_dev, A, B, X, prec, tol, max_iters, gmres, terms,
                 pivoting, x_ref)


def posv_mpir(A: DeviceMatrix, B: DeviceMatrix, X: DeviceMatrix, prec=PREC_FP16, tol=1e-14,
              max_iters=50, gmres=0, terms=0, x_ref=None) -> MpirInfo:
    """Mixed-precision Cholesky + iterative refinement (no reference counterpart)."""
    return _mpir(_lib().ddla_dposv_mpir_dev, A, B, X, prec, tol, max_iters, gmres, terms,
                 PIVOT_PARTIAL, x_ref)


def sgetrf_mp(A, prec=PREC_FP32, pivoting=PIVOT_TOURNAMENT):
    """The mixed solver's FP32 LU factor on its own (no reference counterpart).

    A: square float32 host array (not modified).  Returns (LU, perm, info): unit L below the
    diagonal and U on and above it, row i of LU belonging to row perm[i] of A; info > 0 is the
    first zero pivot column + 1."""
    a = np.ascontiguousarray(A, dtype=np.float32)
    n = a.shape[0]
    assert a.shape == (n, n)
    buf, perm = _DevBuf(a.nbytes), _DevBuf(4 * n)
    buf.upload(a)
    st = _lib().ddla_sgetrf_mp_dev(n, buf.ptr, n, prec, pivoting, perm.ptr)
    if st < 0:
        check(st)
    return (buf.download(np.float32, n * n).reshape(n, n), perm.download(np.int32, n), st)


def sgetf2(P):
    """The FP32 panel of the mixed LU on its own (PLUDecUnblock, LU.cs:344-374, in FP32; no
    reference counterpart in this precision).

    P: m x nb float32 host array (not modified), 1 <= nb <= min(m, 128).  Returns (LU, piv,
    info): the factored panel, relative pivots (nb entries) and the first zero pivot column + 1
    (0 if none)."""
    a = np.ascontiguousarray(P, dtype=np.float32)
    m, nb = a.shape
    buf, piv = _DevBuf(a.nbytes), _DevBuf(4 * nb)
    buf.upload(a)
    st = _lib().ddla_sgetf2_dev(m, nb, buf.ptr, nb, piv.ptr)
    if st < 0:
        check(st)
    return buf.download(np.float32, m * nb).reshape(m, nb), piv.download(np.int32, nb), st


# ------------------------------------------------------------ batched factorizations ---
BATCH_LU, BATCH_CHOL = 0, 1


class _DevBuf:
    """Raw device buffer from ddla_malloc (freed with the object)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(_lib().ddla_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(_lib().ddla_memcpy_h2d(self.ptr, _ptr(arr), arr.nbytes))

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        check(_lib().ddla_memcpy_d2h(_ptr(out), self.ptr, out.nbytes))
        return out

    def __del__(self):
        try:
            if self.ptr and not native.finalized():
                _lib().ddla_free(self.ptr)
        except Exception:
            pass


class BatchPlan:
    """Batched factorization of many small problems of uneven order (no reference
    counterpart; each problem is what LU.PLUDecompose / Cholesky.CholDecompose do for one
    matrix).  Problems are packed row-major one after another in device memory; the plan
    schedules them over one persistent launch captured in a hipGraph (ddla_batch_create).

    a_ptr / piv_ptr / info_ptr are device addresses (ints or c_void_p) of the packed matrices,
    pivots (sum n) and per-problem status (count).  tune: 0 (or False) a fixed shape, 1 (or
    True) the tuning cache's shape for the batch's bucket or a timed choice that is then
    stored, 2 always timed and stored."""

    def __init__(self, n, kind, prec, a_ptr, piv_ptr, info_ptr, tune=False):
        self.n = np.ascontiguousarray(n, dtype=np.int32)
        self.kind = np.ascontiguousarray(kind, dtype=np.int32)
        assert self.n.shape == self.kind.shape
        self.prec = prec
        h = ctypes.c_void_p()
        check(_lib().ddla_batch_create(ctypes.byref(h), prec, self.n.size, _ptr(self.n),
                                       _ptr(self.kind), ctypes.c_void_p(int(_addr(a_ptr))),
                                       ctypes.c_void_p(int(_addr(piv_ptr))),
                                       ctypes.c_void_p(int(_addr(info_ptr))), int(tune)))
        self.handle = h

    def execute(self):
        check(_lib().ddla_batch_execute(self.handle))

    def describe(self) -> dict:
        import json
        buf = ctypes.create_string_buffer(4096)
        check(_lib().ddla_batch_describe(self.handle, buf, len(buf)))
        return json.loads(buf.value.decode())

    def __del__(self):
        try:
            if getattr(self, "handle", None) and not native.finalized():
                _lib().ddla_batch_destroy(self.handle)
        except Exception:
            pass


def _addr(p):
    if isinstance(p, ctypes.c_void_p):
        return p.value or 0
    return int(p)


def batch_factor(mats, kinds, prec=PREC_FP64, tune=False):
    """Factors host matrices (list of square numpy arrays, order <= 256) in one batched launch.
    kinds[i] is BATCH_LU or BATCH_CHOL.  Returns (factors, pivots, info): factors in the
    layout of PLUDecompose / CholDecompose(LOWER) (the upper triangle of a Cholesky problem
    keeps its input), relative pivots per LU problem (None for Cholesky), status per problem."""
    dt = np.float64 if prec == PREC_FP64 else np.float32
    n = np.array([m.shape[0] for m in mats], dtype=np.int32)
    kinds = np.asarray(kinds, dtype=np.int32)
    flat = np.concatenate([np.ascontiguousarray(m, dtype=dt).ravel() for m in mats]) if len(mats) \
        else np.zeros(0, dt)
    a = _DevBuf(flat.nbytes)
    a.upload(flat)
    piv = _DevBuf(int(n.sum()) * 4)
    info = _DevBuf(len(mats) * 4)
    plan = BatchPlan(n, kinds, prec, a.ptr, piv.ptr, info.ptr, tune=tune)
    plan.execute()
    synchronize()
    out = a.download(dt, flat.size)
    pv = piv.download(np.int32, int(n.sum()))
    st = info.download(np.int32, len(mats))
    del plan
    factors, pivots = [], []
    o = po = 0
    for i, ni in enumerate(n):
        factors.append(out[o:o + ni * ni].reshape(ni, ni).copy())
        pivots.append(pv[po:po + ni].copy() if kinds[i] == BATCH_LU else None)
        o += ni * ni
        po += ni
    return factors, pivots, st
