// The factorizations of DDLA (LU.cs, Cholesky.cs, LDLT.cs, QR.cs) on MatrixView operands,
// retargeted to libddla_hip.so: the same static entry points, argument meaning and exceptions
// as the managed code they replace, with the work on the GPU (host views are copied to the
// device and back inside each call).  Build inside the DDLA project with DDLA_HIP defined;
// LU, Cholesky, LDLT and QR then call these in place of their managed blocked loops.
#if DDLA_HIP
using System;
using System.Runtime.InteropServices;
using DDLA.Core;
using DDLA.Misc.Flags;

namespace DDLA.Hip;

public static class HipFactorizations
{
    private const string Lib = HipNative.Lib;

    [DllImport(Lib, EntryPoint = "ddla_dgetrf")]
    private static extern int dgetrf(long n, ref double a, long rsa, long csa, ref int piv);

    [DllImport(Lib, EntryPoint = "ddla_dgetrs")]
    private static extern int dgetrs(int trans, long n, long nrhs, ref double a, long rsa,
                                     long csa, ref int piv, ref double b, long rsb, long csb);

    [DllImport(Lib, EntryPoint = "ddla_dpotrf")]
    private static extern int dpotrf(int uplo, long n, ref double a, long rsa, long csa);

    [DllImport(Lib, EntryPoint = "ddla_dpotrs")]
    private static extern int dpotrs(int uplo, long n, long nrhs, ref double a, long rsa,
                                     long csa, ref double b, long rsb, long csb);

    [DllImport(Lib, EntryPoint = "ddla_dldlt")]
    private static extern int dldlt(int uplo, long n, ref double a, long rsa, long csa);

    [DllImport(Lib, EntryPoint = "ddla_dldlt_solve")]
    private static extern int dldlt_solve(int uplo, long n, long nrhs, ref double a, long rsa,
                                          long csa, ref double b, long rsb, long csb);

    [DllImport(Lib, EntryPoint = "ddla_dgeqrf")]
    private static extern int dgeqrf(long m, long n, ref double a, long rsa, long csa,
                                     ref double t, long rst, long cst);

    [DllImport(Lib, EntryPoint = "ddla_dqr_solve")]
    private static extern int dqr_solve(long m, long n, long nrhs, ref double a, long rsa,
                                        long csa, ref double t, long rst, long cst, ref double b,
                                        long rsb, long csb, ref double x, long rsx, long csx);

    private static void Square(MatrixView A, string what)
    {
        if (A.Rows != A.Cols)
            throw new ArgumentException($"Matrix must be square for {what}.");
    }

    /// <summary>LU.PLUDecompose (LU.cs:251-264): PA = LU in place, relative pivots; a zero
    /// pivot throws ArgumentOutOfRangeException as the reference's InvScal does.</summary>
    public static void PLUDecompose(MatrixView A, Span<int> pivs)
    {
        Square(A, "LU factorization");
        if (pivs.Length != A.Rows)
            throw new ArgumentException("The length of pivs must match the number of rows in the matrix.");
        if (A.Rows == 0) return;
        int info = HipStatus.Check(dgetrf(A.Rows, ref A.GetHeadRef(), A.RowStride, A.ColStride,
                                          ref pivs[0]));
        if (info > 0)
            throw new ArgumentOutOfRangeException("alpha", $"zero pivot at {info - 1}");
    }

    /// <summary>LU.PLUSolve (LU.cs:266-287, :431-465): B <- A^-1 B (or A^-T B).</summary>
    public static void PLUSolve(MatrixView A, Span<int> pivs, MatrixView B, bool trans = false)
    {
        Square(A, "LU solve");
        if (B.Rows != A.Rows)
            throw new ArgumentException("Matrix dimensions do not match for LU solve.");
        if (A.Rows == 0 || B.Cols == 0) return;
        HipStatus.Check(dgetrs(trans ? HipNative.Transpose : HipNative.NoTranspose, A.Rows,
                               B.Cols, ref A.GetHeadRef(), A.RowStride, A.ColStride, ref pivs[0],
                               ref B.GetHeadRef(), B.RowStride, B.ColStride));
    }

    /// <summary>Cholesky.CholDecompose (Cholesky.cs:52-69); a non-positive pivot throws
    /// LinalgException like the lower reference path.</summary>
    public static void CholDecompose(UpLo uplo, MatrixView A)
    {
        Square(A, "Cholesky factorization");
        if (A.Rows == 0) return;
        int info = HipStatus.Check(dpotrf((int)uplo, A.Rows, ref A.GetHeadRef(), A.RowStride,
                                          A.ColStride));
        if (info > 0)
            throw new LinalgException("Cholesky",
                $"Cholesky factorization encountered a non-positive pivot at index {info - 1}.");
    }

    /// <summary>Cholesky.CholeskySolve (Cholesky.cs:102-123, :210-226).</summary>
    public static void CholeskySolve(UpLo uplo, MatrixView A, MatrixView B)
    {
        Square(A, "Cholesky solve");
        if (B.Rows != A.Rows)
            throw new ArgumentException("Matrix dimensions do not match for Cholesky solve.");
        if (A.Rows == 0 || B.Cols == 0) return;
        HipStatus.Check(dpotrs((int)uplo, A.Rows, B.Cols, ref A.GetHeadRef(), A.RowStride,
                               A.ColStride, ref B.GetHeadRef(), B.RowStride, B.ColStride));
    }

    /// <summary>LDLT.LDLTDecompose (LDLT.cs:56-73); a zero pivot throws LinalgException.</summary>
    public static void LDLTDecompose(UpLo uplo, MatrixView A)
    {
        Square(A, "LDLT factorization");
        if (A.Rows == 0) return;
        int info = HipStatus.Check(dldlt((int)uplo, A.Rows, ref A.GetHeadRef(), A.RowStride,
                                         A.ColStride));
        if (info > 0)
            throw new LinalgException("Cholesky",
                $"LDLT factorization encountered a zero pivot at index {info - 1}.");
    }

    /// <summary>LDLT.LDLTSolve (LDLT.cs:105-114, :267-295).</summary>
    public static void LDLTSolve(UpLo uplo, MatrixView A, MatrixView B)
    {
        Square(A, "LDLT solve");
        if (B.Rows != A.Rows)
            throw new ArgumentException("Matrix dimensions do not match for Cholesky solve.");
        if (A.Rows == 0 || B.Cols == 0) return;
        HipStatus.Check(dldlt_solve((int)uplo, A.Rows, B.Cols, ref A.GetHeadRef(), A.RowStride,
                                    A.ColStride, ref B.GetHeadRef(), B.RowStride, B.ColStride));
    }

    /// <summary>QR.QRDecompose(A, T) (QR.cs:138-149): T has min(256, max(m, n)) rows and
    /// A.Cols columns (QR.CreateT).</summary>
    public static void QRDecompose(MatrixView A, MatrixView T)
    {
        if (T.Cols != A.Cols)
            throw new ArgumentException("Matrix T and A must have the same number of columns.");
        if (T.Rows < Math.Min(256, Math.Max(A.Rows, A.Cols)))
            throw new ArgumentException("Matrix T has too few rows.");
        if (A.Rows == 0 || A.Cols == 0) return;
        HipStatus.Check(dgeqrf(A.Rows, A.Cols, ref A.GetHeadRef(), A.RowStride, A.ColStride,
                               ref T.GetHeadRef(), T.RowStride, T.ColStride));
    }

    /// <summary>QR.Solve(A, T, B, X) (QR.cs:151-167): least squares, m >= n; B kept.</summary>
    public static void QRSolve(MatrixView A, MatrixView T, MatrixView B, MatrixView X)
    {
        if (A.Cols != T.Cols)
            throw new ArgumentException("Matrix T and A must have the same number of columns.");
        if (A.Rows != B.Rows)
            throw new ArgumentException("Matrix A and B must have the same number of rows.");
        if (A.Cols != X.Rows)
            throw new ArgumentException("The number of columns in A must match the number of rows in X.");
        if (X.Cols != B.Cols)
            throw new ArgumentException("Matrix X and B must have the same number of columns.");
        if (A.Cols == 0 || B.Cols == 0) return;
        HipStatus.Check(dqr_solve(A.Rows, A.Cols, B.Cols, ref A.GetHeadRef(), A.RowStride,
                                  A.ColStride, ref T.GetHeadRef(), T.RowStride, T.ColStride,
                                  ref B.GetHeadRef(), B.RowStride, B.ColStride,
                                  ref X.GetHeadRef(), X.RowStride, X.ColStride));
    }
}
#endif
