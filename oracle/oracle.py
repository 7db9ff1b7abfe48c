"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: a CPU restatement of DDLA's managed FP64 GEMM, blocked LU with
partial pivoting, blocked Cholesky and (in numpy) blocked LDL^T and Householder QR, used to check the HIP library and as the timed CPU
baseline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import it.
Every function mirrors the reference routine named in its docstring (paths relative to
the DDLA repository root); the C source cites the same lines.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

NO_TRANS, TRANS = 0x00, 0x08
LOWER, UPPER = 0xC0, 0x60
NON_UNIT, UNIT = 0x000, 0x100
LEFT, RIGHT = 0, 1

_lib = None


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i64, dp, ip = ctypes.c_int64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)
        d, i = ctypes.c_double, ctypes.c_int
        L.orc_set_threads.argtypes = [i]
        L.orc_set_threads.restype = i
        L.orc_get_threads.restype = i
        L.orc_set_block.argtypes = [i, i]
        L.orc_fill_uniform.argtypes = [dp, i64, ctypes.c_uint64]
        L.orc_dgemm.argtypes = [i, i, i64, i64, i64, d, dp, i64, i64, dp, i64, i64, d, dp, i64, i64]
        L.orc_damaxv.argtypes = [i64, dp, i64]
        L.orc_damaxv.restype = i64
        L.orc_dtrsm.argtypes = [i, i, i, i, i64, i64, d, dp, i64, i64, dp, i64, i64]
        L.orc_dsyrk.argtypes = [i, i, i64, i64, d, dp, i64, i64, d, dp, i64, i64]
        L.orc_dgetrf.argtypes = [i64, dp, i64, i64, ip]
        L.orc_dgetrf.restype = i64
        L.orc_dgetf2.argtypes = [i64, i64, dp, i64, i64, ip]
        L.orc_dgetf2.restype = i64
        fp = ctypes.POINTER(ctypes.c_float)
        L.orc_stslu.argtypes = [i64, fp, i64, ip, fp, ip, ip]
        L.orc_sgetf2.argtypes = [i64, i64, fp, i64, i64, ip]
        L.orc_sgetf2.restype = i64
        L.orc_stslu.restype = i
        L.orc_dgetrf_nopiv.argtypes = [i64, dp, i64, i64]
        L.orc_dgetrf_nopiv.restype = i64
        L.orc_dlaswp.argtypes = [i, i, i64, i64, dp, i64, i64, ip, i64]
        L.orc_dgetrs.argtypes = [i, i64, i64, dp, i64, i64, ip, dp, i64, i64]
        L.orc_dpotrf.argtypes = [i, i64, dp, i64, i64]
        L.orc_dpotrf.restype = i64
        L.orc_dpotrs.argtypes = [i, i64, i64, dp, i64, i64, dp, i64, i64]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _strides(a):
    return a.strides[0] // 8, a.strides[1] // 8


def set_threads(t: int) -> int:
    return lib().orc_set_threads(int(t))


def threads() -> int:
    return lib().orc_get_threads()


def set_block(piv_block: int = 128, chol_block: int = 128):
    lib().orc_set_block(int(piv_block), int(chol_block))


def uniform(n_elems: int, seed: int) -> np.ndarray:
    """SplitMix64 counter stream mapped to [-1, 1): x_i = (z_i >> 11) * 2^-52 - 1."""
    out = np.empty(int(n_elems), dtype=np.float64)
    lib().orc_fill_uniform(_dp(out), out.size, ctypes.c_uint64(seed & (2**64 - 1)))
    return out


def uniform_matrix(m: int, n: int, seed: int) -> np.ndarray:
    return uniform(m * n, seed).reshape(m, n)


def gemm(alpha, A, B, beta, C, transa=NO_TRANS, transb=NO_TRANS):
    """Managed BlasProvider.GeMM (DDLA/BLAS/Managed/3m.cs:13-49); C updated in place."""
    assert A.dtype == B.dtype == C.dtype == np.float64
    m, n = C.shape
    k = A.shape[0] if transa == TRANS else A.shape[1]
    lib().orc_dgemm(transa, transb, m, n, k, alpha, _dp(A), *_strides(A), _dp(B), *_strides(B),
                    beta, _dp(C), *_strides(C))
    return C


def matmul(A, B):
    """C = A * B through the managed GEMM (MatrixView.operator*, Core/MatrixView.cs:844-854)."""
    C = np.empty((A.shape[0], B.shape[1]), dtype=np.float64)
    return gemm(1.0, A, B, 0.0, C)


def amax(x):
    """AMax (DDLA/BLAS/1v.cs:30-41)."""
    x = np.asarray(x, dtype=np.float64)
    return int(lib().orc_damaxv(x.size, _dp(x), x.strides[0] // 8))


def trsm(side, uplo, trans, diag, alpha, A, B):
    """TRSM (facade DDLA/BLAS/3m.cs:287-314); B overwritten with the solution."""
    m, n = B.shape
    lib().orc_dtrsm(side, uplo, trans, diag, m, n, alpha, _dp(A), *_strides(A), _dp(B), *_strides(B))
    return B


def syrk(uplo, trans, alpha, A, beta, C):
    """SYRK on the uplo triangle (facade DDLA/BLAS/3m.cs:118-147)."""
    m = C.shape[0]
    k = A.shape[0] if trans == TRANS else A.shape[1]
    lib().orc_dsyrk(uplo, trans, m, k, alpha, _dp(A), *_strides(A), beta, _dp(C), *_strides(C))
    return C


def _op(X, trans):
    return X.T if trans == TRANS else X


def _expand(A, uplo, sym=False, unit=False):
    """The dense matrix a BLIS routine reads from the uplo triangle of A: symmetric (sym) or
    triangular with zeros outside the triangle (ones on the diagonal when unit)."""
    A = np.asarray(A, dtype=np.float64)
    tri = np.tril(A) if uplo == LOWER else np.triu(A)
    if sym:
        return tri + tri.T - np.diag(np.diag(A))
    if unit:
        tri = tri - np.diag(np.diag(tri)) + np.eye(A.shape[0])
    return tri


def _keep_triangle(C, full, uplo):
    mask = np.tril(np.ones(C.shape, bool)) if uplo == LOWER else np.triu(np.ones(C.shape, bool))
    C[mask] = full[mask]
    return C


def gemmt(uplo, transa, transb, alpha, A, B, beta, C):
    """GeMMT (facade DDLA/BLAS/3m.cs:44-70 -> bli_dgemmt; managed DDLA/BLAS/Managed/3m.cs:51-131):
    the uplo triangle of C = beta*C + alpha*op(A)*op(B), the other triangle untouched.  The
    triangle takes the restated managed GEMM's sums."""
    full = gemm(alpha, np.ascontiguousarray(A), np.ascontiguousarray(B), beta,
                np.ascontiguousarray(C, dtype=np.float64).copy(), transa, transb)
    return _keep_triangle(C, full, uplo)


def syr2k(uplo, transa, transb, alpha, A, B, beta, C):
    """SyR2k (facade DDLA/BLAS/3m.cs:149-172; managed Managed/3m.cs:227-245: two GeMMT calls,
    the second with beta = 1)."""
    gemmt(uplo, transa, NO_TRANS, alpha, A, np.ascontiguousarray(_op(B, transb).T), beta, C)
    return gemmt(uplo, transb, NO_TRANS, alpha, B, np.ascontiguousarray(_op(A, transa).T), 1.0, C)


def symm(side, uplo, transb, alpha, A, B, beta, C):
    """SyMM with BLIS's semantics (facade DDLA/BLAS/3m.cs:83-106 -> bli_dsymm): C = beta*C +
    alpha*S*op(B) (LEFT) or alpha*op(B)*S (RIGHT), S symmetric from the uplo triangle of A."""
    S = _expand(A, uplo, sym=True)
    Bo = np.ascontiguousarray(_op(B, transb))
    if side == LEFT:
        return gemm(alpha, S, Bo, beta, C)
    return gemm(alpha, Bo, S, beta, C)


def trmm3(side, uplo, transa, diag, transb, alpha, A, B, beta, C):
    """TrMM3 with BLIS's semantics (facade DDLA/BLAS/3m.cs:225-250 -> bli_dtrmm3): C = beta*C +
    alpha*op(T)*op(B) (LEFT) or alpha*op(B)*op(T) (RIGHT), T the uplo triangle of A."""
    T = np.ascontiguousarray(_op(_expand(A, uplo, unit=diag == UNIT), transa))
    Bo = np.ascontiguousarray(_op(B, transb))
    if side == LEFT:
        return gemm(alpha, T, Bo, beta, C)
    return gemm(alpha, Bo, T, beta, C)


def trmm(side, uplo, trans, diag, alpha, A, B):
    """TrMM with BLIS's semantics (facade DDLA/BLAS/3m.cs:185-204 -> bli_dtrmm): B = alpha*op(T)*B
    (LEFT) or alpha*B*op(T) (RIGHT).  The managed TrMM (Managed/3m.cs:259-287) does not apply
    aTrans; the drop-in follows BLIS, which does."""
    out = np.zeros(B.shape)
    trmm3(side, uplo, trans, diag, NO_TRANS, alpha, A, np.ascontiguousarray(B), 0.0, out)
    B[...] = out
    return B


def getrf(A):
    """PLUDecompose (LU.cs:251-264) in place. Returns (info, relative pivots int32[n])."""
    n = A.shape[0]
    piv = np.zeros(n, dtype=np.int32)
    info = lib().orc_dgetrf(n, _dp(A), *_strides(A), _ip(piv))
    return int(info), piv


def getf2(P):
    """PLUDecUnblock (LU.cs:344-374) on an m x nb panel in place."""
    m, nb = P.shape
    piv = np.zeros(min(m, nb), dtype=np.int32)
    info = lib().orc_dgetf2(m, nb, _dp(P), *_strides(P), _ip(piv))
    return int(info), piv


def sgetf2(P):
    """PLUDecUnblock (LU.cs:344-374) in FP32 on an m x nb float32 panel in place (the mixed
    LU's panel; the reference has no FP32 path).  Returns (info, relative pivots)."""
    assert P.dtype == np.float32
    m, nb = P.shape
    piv = np.zeros(min(m, nb), dtype=np.int32)
    fp = ctypes.POINTER(ctypes.c_float)
    info = lib().orc_sgetf2(m, nb, P.ctypes.data_as(fp), P.strides[0] // 4, P.strides[1] // 4,
                            _ip(piv))
    return int(info), piv


def tslu(P):
    """Tournament-pivoting panel (FP32; no reference counterpart, CALU's tournament over stacks
    of 256 rows).  P: m x 128 float32, m >= 128 (not modified).  Returns (zero, sel, top,
    pairs): zero = first zero-pivot column + 1 or 0, sel = panel rows in pivot order, top =
    their factored rows (L11 \\ U11), pairs = (dst, src) row moves."""
    P = np.ascontiguousarray(P, dtype=np.float32)
    m, nb = P.shape
    assert nb == 128 and m >= 128
    sel = np.zeros(128, dtype=np.int32)
    top = np.zeros((128, 128), dtype=np.float32)
    lst = np.zeros(512, dtype=np.int32)
    cnt = np.zeros(1, dtype=np.int32)
    fp = ctypes.POINTER(ctypes.c_float)
    zero = lib().orc_stslu(m, P.ctypes.data_as(fp), nb, _ip(sel), top.ctypes.data_as(fp),
                           _ip(lst), _ip(cnt))
    return int(zero), sel, top, lst[: 2 * int(cnt[0])].reshape(-1, 2)


def getrf_nopiv(A):
    """LUDecompose (LU.cs:133-142) in place."""
    return int(lib().orc_dgetrf_nopiv(A.shape[0], _dp(A), *_strides(A)))


def laswp(X, piv, side=LEFT, forward=True):
    """ApplyPiv (LU.cs:237-249) in place."""
    piv = np.ascontiguousarray(piv, dtype=np.int32)
    m, n = X.shape
    lib().orc_dlaswp(side, 1 if forward else 0, m, n, _dp(X), *_strides(X), _ip(piv), piv.size)
    return X


def getrs(A, piv, B, trans=False):
    """PLUSolve (LU.cs:266-287,431-465); B overwritten with X."""
    piv = np.ascontiguousarray(piv, dtype=np.int32)
    n, nrhs = B.shape
    lib().orc_dgetrs(TRANS if trans else NO_TRANS, n, nrhs, _dp(A), *_strides(A), _ip(piv),
                     _dp(B), *_strides(B))
    return B


def potrf(A, uplo=LOWER):
    """CholDecompose (Cholesky.cs:52-69) in place; returns info (0 or first bad pivot+1)."""
    return int(lib().orc_dpotrf(uplo, A.shape[0], _dp(A), *_strides(A)))


def potrs(A, B, uplo=LOWER):
    """CholeskySolve (Cholesky.cs:102-123,210-226); B overwritten with X."""
    n, nrhs = B.shape
    lib().orc_dpotrs(uplo, n, nrhs, _dp(A), *_strides(A), _dp(B), *_strides(B))
    return B


def _ldlt_unblocked(At):
    """LDLTLowerUnblock (DDLA/Factorizations/LDLT.cs:141-158) on the lower triangle of At, in
    place: for i < n - 1 a zero a_ii stops with i + 1, else a21 *= 1/a_ii (InvScaled) and the
    lower triangle of A22 gets a_rc += (-a_ii l_r) l_c.  The last pivot is not tested (as in
    the reference)."""
    n = At.shape[0]
    for i in range(n - 1):
        d = At[i, i]
        if d == 0.0:
            return i + 1
        At[i + 1:, i] *= 1.0 / d
        l = At[i + 1:, i].copy()
        for r in range(i + 1, n):
            At[r, i + 1:r + 1] += (-d * l[r - i - 1]) * l[:r - i]
    return 0


def ldlt(A, uplo=LOWER, block=128):
    """LDLTDecompose (LDLT.cs:56-73) in place; returns info (0, or k + 1 for the first zero
    pivot the reference tests).  Lower: LDLTLowerBlock (:160-202): per block, the unblocked
    factor of A11, A21 <- A21 L11^-T (unit), Y21 = A21, A21[:, j] *= 1/d_j, lower(A22) -=
    Y21 A21^T.  Upper (LDLTUpperBlock, :223-265) is the same sequence on the transposed
    storage (A = U^T D U), which is how it is run here.  Unblocked when n <= block."""
    import scipy.linalg as sla
    At = A if uplo == LOWER else A.T
    n = At.shape[0]
    if n <= block:
        return _ldlt_unblocked(At)
    for j in range(0, n, block):
        b = min(block, n - j)
        info = _ldlt_unblocked(At[j:j + b, j:j + b])
        if info:
            return j + info
        if j + b >= n:
            break
        A11 = At[j:j + b, j:j + b]
        A21 = At[j + b:, j:j + b]
        A21[...] = sla.solve_triangular(A11, A21.T, lower=True, unit_diagonal=True).T
        Y = A21.copy()
        A21 *= 1.0 / np.diag(A11)[None, :]
        A22 = At[j + b:, j + b:]
        A22[...] = np.where(np.tril(np.ones(A22.shape, bool)), A22 - Y @ A21.T, A22)
    return 0


def ldlt_solve(A, B, uplo=LOWER):
    """LDLTSolveUncheck (LDLT.cs:267-295): X <- L^-1 X, X[i, :] *= 1/d_i, X <- L^-T X with
    L the unit lower triangle (U^T for UPPER); B overwritten with X."""
    import scipy.linalg as sla
    At = A if uplo == LOWER else A.T
    X = sla.solve_triangular(At, B, lower=True, unit_diagonal=True)
    X *= (1.0 / np.diag(At))[:, None]
    B[...] = sla.solve_triangular(At.T, X, lower=False, unit_diagonal=True)
    return B


def ldlt_factors(F, uplo=LOWER):
    """(T, d): the unit triangle of the chosen side and the diagonal of an LDL^T factor
    (LDLT.Deconstruct, LDLT.cs:128-137)."""
    T = np.tril(F, -1) + np.eye(F.shape[0]) if uplo == LOWER else np.triu(F, 1) + np.eye(F.shape[0])
    return T, np.diag(F).copy()


QR_BLOCK = 256  # QR.BlockSize (DDLA/Factorizations/QR.cs:16)


def qr_t_rows(m, n):
    """Rows of the T matrix QR.CreateT allocates (QR.cs:124-128): min(256, max(m, n))."""
    return min(QR_BLOCK, max(m, n))


def _qr_unblocked(A, T):
    """QRDecUnblock (QR.cs:326-363) on the view A (m x n), T (>= min(m, n) rows): per column
    j, BuildHouseHolder (Transformations/HouseHolder.cs:34-56: rho = -copysign(hypot(chi,
    |x2|), chi), u2 = x2 * (1/(chi - rho)), tau = (1 + |u2|^2 / (chi - rho)^2) / 2; tau = 1/2
    and chi <- -chi when x2 = 0), ApplyHouseHolder from the left (:95-111: w = (a12 + A22^T
    u2) * (1/tau), a12 -= w, A22 -= u2 w^T), T[j, j] = tau and T[:j, j] = a10 + A20^T u2.
    The norm is sqrt of the sum of squares (the reference calls BLIS's scaled norm)."""
    m, n = A.shape
    for j in range(min(m, n)):
        chi = A[j, j]
        x2 = A[j + 1:, j]
        len2 = float(np.sqrt(np.dot(x2, x2)))
        if len2 == 0.0:
            A[j, j] = -chi
            tau = 0.5
        else:
            rho = -np.copysign(np.hypot(chi, len2), chi)
            miu = chi - rho
            A[j, j] = rho
            x2 *= 1.0 / miu
            s = len2 / abs(miu)
            tau = (1.0 + s * s) / 2.0
        u2 = A[j + 1:, j]
        if j + 1 < n:
            w = A[j, j + 1:] + u2 @ A[j + 1:, j + 1:]
            w *= 1.0 / tau
            A[j, j + 1:] -= w
            A[j + 1:, j + 1:] -= np.outer(u2, w)
        T[j, j] = tau
        T[:j, j] = A[j, :j] + A[j + 1:, :j].T @ u2
    return A, T


def _unit_lower(P):
    """The Householder vectors of a factored panel P (m x b): unit diagonal, zeros above."""
    U = np.tril(P, -1)
    k = min(P.shape)
    U[np.arange(k), np.arange(k)] = 1.0
    return U


def _qr_blocks(m, n, tb):
    k, out = 0, []
    while k < min(m, n):
        b = min(tb, min(m, n) - k)
        out.append((k, b))
        k += b
    return out


def qr(A, T=None):
    """QRDecompose (QR.cs:130-149) in place: unblocked when n <= 256, else QRDecBlock
    (:365-395): per block of b = min(T rows, remaining) columns the unblocked factor of the
    panel A[k:, k:k+b] (its T factor in T[:b, k:k+b]), then ApplyQlhfc of the panel to
    A[k:, k+b:].  Returns T (qr_t_rows(m, n) x n; entries outside the factors' upper
    triangles are unspecified)."""
    m, n = A.shape
    tb = qr_t_rows(m, n)
    if T is None:
        T = np.zeros((tb, n))
    if n <= QR_BLOCK:
        _qr_unblocked(A, T)
        return T
    for k, b in _qr_blocks(m, n, tb):
        _qr_unblocked(A[k:, k:k + b], T[:b, k:k + b])
        if k + b < n:
            _apply_block(A[k:, k:k + b], T[:b, k:k + b], A[k:, k + b:], trans=True)
    return T


def _apply_block(P, Tb, B, trans):
    """B <- (I - U Tb^-1 U^T)^T B (trans) or (I - U Tb^-1 U^T) B, U the unit lower part of the
    panel P (ApplyQlhfc / ApplyQlnfc's step, QR.cs:444-463, :505-519)."""
    import scipy.linalg as sla
    U = _unit_lower(P)
    W = U.T @ B
    W = sla.solve_triangular(np.triu(Tb), W, lower=False, trans=1 if trans else 0)
    B -= U @ W


def qr_apply(A, T, B, trans):
    """ApplyQlhfc (trans: B <- Q^T B, QR.cs:423-464) or ApplyQlnfc (B <- Q B, :466-521) with
    the blocks of the factorization (forward for Q^T, backward for Q)."""
    m, n = A.shape
    blocks = _qr_blocks(m, n, T.shape[0])
    for k, b in (blocks if trans else blocks[::-1]):
        _apply_block(A[k:, k:k + b], T[:b, k:k + b], B[k:], trans)
    return B


def qr_solve(A, T, B):
    """SolveUncheck (QR.cs:397-409): Y = Q^T B, X = R^-1 Y[:n] (least squares, m >= n)."""
    import scipy.linalg as sla
    n = A.shape[1]
    Y = qr_apply(A, T, np.array(B, dtype=np.float64, copy=True), True)
    return sla.solve_triangular(np.triu(A[:n, :n]), Y[:n], lower=False)


def contract(expression, alpha, A, B, beta, C):
    """C = beta C + alpha contract(expression, A, B) with MArray.Contract's index rules
    (DDLA/Einsum/MArray.Contract.cs:35-125; IndiceUtils.Diagonal / Divide,
    Utilities/IndiceUtils.cs:18-130), restated as explicit loops over the index classes:
    a repeated symbol within one tensor is its diagonal; symbols of A and B only are summed
    products (K), of A and C the rows (M), of B and C the columns (N), of all three a batch;
    symbols of A alone or B alone are summed, of C alone broadcast.  C is updated in place
    (every element scaled by beta first, then alpha = 0 returns)."""
    arrow = expression.split("->")
    if len(arrow) != 2:
        raise ValueError("Expression must contain exactly one '->' symbol.")
    ops = arrow[0].split(",")
    if len(ops) != 2:
        raise ValueError("Expression must contain exactly two symbols before '->'.")
    sa, sb, sc = ops[0].strip(), ops[1].strip(), arrow[1].strip()

    def diag(sym, T):
        if len(sym) != T.ndim:
            raise ValueError("The rank of the tensor must match the length of the symbol.")
        out = {}
        for i, ch in enumerate(sym):
            L, st = T.shape[i], T.strides[i] // T.itemsize
            if ch in out:
                if out[ch][0] != L:
                    raise ValueError(f"The character '{ch}' already exists with different lengths.")
                out[ch] = (L, out[ch][1] + st)
            else:
                out[ch] = (L, st)
        return out

    ia, ib, ic = diag(sa, A), diag(sb, B), diag(sc, C)
    for ch in set(ia) | set(ib) | set(ic):
        ls = {d[ch][0] for d in (ia, ib, ic) if ch in d}
        if len(ls) > 1:
            raise ValueError(f"The character '{ch}' exists in both indices with different lengths.")
    C[...] = 0.0 if beta == 0.0 else beta * C
    if alpha == 0.0:
        return C
    syms = sorted(set(ia) | set(ib) | set(ic))
    lens = [next(d[ch][0] for d in (ia, ib, ic) if ch in d) for ch in syms]
    # a flat view over the memory each tensor spans (non-negative strides), indexed by the
    # element strides of the merged indices
    def flat(T):
        st = np.array(T.strides, dtype=np.int64) // T.itemsize
        if (st < 0).any():
            raise ValueError("negative strides are not handled here")
        hi = int(sum((n - 1) * s for n, s in zip(T.shape, st))) if T.size else 0
        return np.lib.stride_tricks.as_strided(T, shape=(hi + 1,), strides=(T.itemsize,)), 0

    (fa_, oa), (fb_, ob), (fc_, oc) = flat(A), flat(B), flat(C)
    if any(L == 0 for L in lens):
        return C
    for multi in np.ndindex(*lens):
        offa = oa + sum(i * ia[ch][1] for i, ch in zip(multi, syms) if ch in ia)
        offb = ob + sum(i * ib[ch][1] for i, ch in zip(multi, syms) if ch in ib)
        offc = oc + sum(i * ic[ch][1] for i, ch in zip(multi, syms) if ch in ic)
        fc_[offc] += alpha * (fa_[offa] * fb_[offb])
    return C


def lu_factors(LU):
    """Split a packed LU into unit-lower L and upper U (LU.Deconstruct, LU.cs:112-128)."""
    L = np.tril(LU, -1) + np.eye(LU.shape[0])
    U = np.triu(LU)
    return L, U


def perm_matrix(piv):
    """P with P*A = L*U, built like LU.Deconstruct: ApplyPiv on the identity."""
    n = len(piv)
    P = np.eye(n)
    laswp(P, piv)
    return P
