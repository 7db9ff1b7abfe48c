// P/Invoke declarations of every entry point of libddla_hip.so (include/ddla_native.h), one
// to one: the same names, argument order and C types (int64 dimensions and strides as long,
// device and host buffers as IntPtr, scalars passed by pointer as `in double`).  Every call
// returns a status: < 0 an error (ddla_last_error holds the message), > 0 a factorization's
// info (1-based index of the failing pivot), 0 success.
//
// The library resolves as "ddla_hip" (libddla_hip.so on Linux).  C# source for a .NET 8
// project; DeviceMatrix.cs needs AllowUnsafeBlocks (it pins host arrays for the copies).
using System;
using System.Runtime.InteropServices;
using System.Text;

namespace DDLA.Hip;

[StructLayout(LayoutKind.Sequential)]
public struct MpirOpts
{
    public int FactorPrec;   // PrecFp16 / PrecBf16 / PrecFp32 / PrecFp64
    public int MaxIters;     // <= 0: 50
    public double Tol;       // <= 0: 1e-14
    public int Gmres;        // must be 0
    public int Terms;        // low-precision parts per FP32 operand (0: default)
    public int Pivoting;     // PivotPartial / PivotButterfly / PivotTournament
    public int Reserved;     // must be 0
    public IntPtr XRef;      // optional device reference solution (n x nrhs)
    public long LdXRef;
}

[StructLayout(LayoutKind.Sequential)]
public struct MpirInfo
{
    public int Status;
    public int Iters;
    public int InnerIters;
    public int Fallback;
    public double BackwardError;
    public double FactorMs;
    public double RefineMs;
    public double ForwardError;  // -1 when no reference solution was given
}

public static class HipNative
{
    public const string Lib = "ddla_hip";

    public const int ErrInvalidArg = -1, ErrHip = -2, ErrNoDevice = -3, ErrAlloc = -4,
        ErrUnsupported = -5, ErrNotConverged = -6;
    public const int NoTranspose = 0x00, Transpose = 0x08;
    public const int Lower = 0xC0, Upper = 0x60;
    public const int NonUnitDiag = 0x000, UnitDiag = 0x100;
    public const int Left = 0, Right = 1;
    public const int PrecFp64 = 0, PrecFp32 = 1, PrecFp16 = 2, PrecBf16 = 3;
    public const int PivotPartial = 0, PivotButterfly = 1, PivotTournament = 2;

    [DllImport(Lib, EntryPoint = "ddla_init")]
    public static extern int ddla_init(int device);

    [DllImport(Lib, EntryPoint = "ddla_finalize")]
    public static extern int ddla_finalize();

    [DllImport(Lib, EntryPoint = "ddla_last_error")]
    public static extern int ddla_last_error(StringBuilder buf, nuint len);

    [DllImport(Lib, EntryPoint = "ddla_set_stream")]
    public static extern int ddla_set_stream(IntPtr stream);

    [DllImport(Lib, EntryPoint = "ddla_get_stream")]
    public static extern int ddla_get_stream(out IntPtr stream);

    [DllImport(Lib, EntryPoint = "ddla_sync")]
    public static extern int ddla_sync();

    [DllImport(Lib, EntryPoint = "ddla_version")]
    public static extern int ddla_version();

    [DllImport(Lib, EntryPoint = "ddla_set_wait_limit")]
    public static extern int ddla_set_wait_limit(ulong rounds);

    [DllImport(Lib, EntryPoint = "ddla_profile_enable")]
    public static extern int ddla_profile_enable(int on);

    [DllImport(Lib, EntryPoint = "ddla_profile_reset")]
    public static extern int ddla_profile_reset();

    [DllImport(Lib, EntryPoint = "ddla_profile_only")]
    public static extern int ddla_profile_only(string name);

    [DllImport(Lib, EntryPoint = "ddla_profile_read")]
    public static extern int ddla_profile_read(int idx, StringBuilder name, nuint len, out long count, out double total_ms, out double flops, out double bytes);

    [DllImport(Lib, EntryPoint = "ddla_malloc")]
    public static extern int ddla_malloc(out IntPtr dptr, nuint bytes);

    [DllImport(Lib, EntryPoint = "ddla_free")]
    public static extern int ddla_free(IntPtr dptr);

    [DllImport(Lib, EntryPoint = "ddla_memcpy_h2d")]
    public static extern int ddla_memcpy_h2d(IntPtr dst_dev, IntPtr src_host, nuint bytes);

    [DllImport(Lib, EntryPoint = "ddla_memcpy_d2h")]
    public static extern int ddla_memcpy_d2h(IntPtr dst_host, IntPtr src_dev, nuint bytes);

    [DllImport(Lib, EntryPoint = "ddla_memcpy_d2d")]
    public static extern int ddla_memcpy_d2d(IntPtr dst_dev, IntPtr src_dev, nuint bytes);

    [DllImport(Lib, EntryPoint = "ddla_dmat_set")]
    public static extern int ddla_dmat_set(long m, long n, IntPtr host, long rs, long cs, IntPtr dev, long ld);

    [DllImport(Lib, EntryPoint = "ddla_dmat_get")]
    public static extern int ddla_dmat_get(long m, long n, IntPtr dev, long ld, IntPtr host, long rs, long cs);

    [DllImport(Lib, EntryPoint = "ddla_drandm_dev")]
    public static extern int ddla_drandm_dev(long m, long n, IntPtr dev, long ld, ulong seed);

    [DllImport(Lib, EntryPoint = "ddla_dgemm")]
    public static extern int ddla_dgemm(int transa, int transb, long m, long n, long k, in double alpha, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb, in double beta, IntPtr c, long rsc, long csc);

    [DllImport(Lib, EntryPoint = "ddla_dgemm_dev")]
    public static extern int ddla_dgemm_dev(int transa, int transb, long m, long n, long k, in double alpha, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb, in double beta, IntPtr c, long rsc, long csc);

    [DllImport(Lib, EntryPoint = "ddla_dgetrf_dev")]
    public static extern int ddla_dgetrf_dev(long n, IntPtr a, long lda, IntPtr dpiv);

    [DllImport(Lib, EntryPoint = "ddla_dgetrf_noalloc_dev")]
    public static extern int ddla_dgetrf_noalloc_dev(long n, IntPtr a, long lda, IntPtr dpiv);

    [DllImport(Lib, EntryPoint = "ddla_dgetrs_dev")]
    public static extern int ddla_dgetrs_dev(int trans, long n, long nrhs, IntPtr a, long lda, IntPtr dpiv, IntPtr b, long ldb);

    [DllImport(Lib, EntryPoint = "ddla_dgetrf")]
    public static extern int ddla_dgetrf(long n, IntPtr a, long rsa, long csa, IntPtr piv);

    [DllImport(Lib, EntryPoint = "ddla_dgetrs")]
    public static extern int ddla_dgetrs(int trans, long n, long nrhs, IntPtr a, long rsa, long csa, IntPtr piv, IntPtr b, long rsb, long csb);

    [DllImport(Lib, EntryPoint = "ddla_dlaswp_dev")]
    public static extern int ddla_dlaswp_dev(int side, int forward, long m, long n, IntPtr x, long ldx, IntPtr dpiv, long npiv);

    [DllImport(Lib, EntryPoint = "ddla_dgetrf_nopiv_dev")]
    public static extern int ddla_dgetrf_nopiv_dev(long n, IntPtr a, long lda);

    [DllImport(Lib, EntryPoint = "ddla_dgetrs_nopiv_dev")]
    public static extern int ddla_dgetrs_nopiv_dev(int trans, long n, long nrhs, IntPtr a, long lda, IntPtr b, long ldb);

    [DllImport(Lib, EntryPoint = "ddla_dpotrf_dev")]
    public static extern int ddla_dpotrf_dev(int uplo, long n, IntPtr a, long lda);

    [DllImport(Lib, EntryPoint = "ddla_dpotrs_dev")]
    public static extern int ddla_dpotrs_dev(int uplo, long n, long nrhs, IntPtr a, long lda, IntPtr b, long ldb);

    [DllImport(Lib, EntryPoint = "ddla_dpotrf")]
    public static extern int ddla_dpotrf(int uplo, long n, IntPtr a, long rsa, long csa);

    [DllImport(Lib, EntryPoint = "ddla_dpotrs")]
    public static extern int ddla_dpotrs(int uplo, long n, long nrhs, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb);

    [DllImport(Lib, EntryPoint = "ddla_dldlt_dev")]
    public static extern int ddla_dldlt_dev(int uplo, long n, IntPtr a, long lda);

    [DllImport(Lib, EntryPoint = "ddla_dldlt_solve_dev")]
    public static extern int ddla_dldlt_solve_dev(int uplo, long n, long nrhs, IntPtr a, long lda, IntPtr b, long ldb);

    [DllImport(Lib, EntryPoint = "ddla_dldlt")]
    public static extern int ddla_dldlt(int uplo, long n, IntPtr a, long rsa, long csa);

    [DllImport(Lib, EntryPoint = "ddla_dldlt_solve")]
    public static extern int ddla_dldlt_solve(int uplo, long n, long nrhs, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb);

    [DllImport(Lib, EntryPoint = "ddla_qr_t_rows")]
    public static extern int ddla_qr_t_rows(long m, long n);

    [DllImport(Lib, EntryPoint = "ddla_dgeqrf_dev")]
    public static extern int ddla_dgeqrf_dev(long m, long n, IntPtr a, long lda, IntPtr t, long ldt);

    [DllImport(Lib, EntryPoint = "ddla_dqr_apply_dev")]
    public static extern int ddla_dqr_apply_dev(int trans, long m, long n, IntPtr a, long lda, IntPtr t, long ldt, long nrhs, IntPtr b, long ldb);

    [DllImport(Lib, EntryPoint = "ddla_dqr_solve_dev")]
    public static extern int ddla_dqr_solve_dev(long m, long n, long nrhs, IntPtr a, long lda, IntPtr t, long ldt, IntPtr b, long ldb, IntPtr x, long ldx);

    [DllImport(Lib, EntryPoint = "ddla_dgeqrf")]
    public static extern int ddla_dgeqrf(long m, long n, IntPtr a, long rsa, long csa, IntPtr t, long rst, long cst);

    [DllImport(Lib, EntryPoint = "ddla_dqr_solve")]
    public static extern int ddla_dqr_solve(long m, long n, long nrhs, IntPtr a, long rsa, long csa, IntPtr t, long rst, long cst, IntPtr b, long rsb, long csb, IntPtr x, long rsx, long csx);

    [DllImport(Lib, EntryPoint = "ddla_dcontract_dev")]
    public static extern int ddla_dcontract_dev(string expr, double alpha, IntPtr a, int na, long[] len_a, long[] str_a, IntPtr b, int nb, long[] len_b, long[] str_b, double beta, IntPtr c, int nc, long[] len_c, long[] str_c);

    [DllImport(Lib, EntryPoint = "ddla_dtrsm_dev")]
    public static extern int ddla_dtrsm_dev(int side, int uplo, int transa, int diag, long m, long n, in double alpha, IntPtr a, long lda, IntPtr b, long ldb);

    [DllImport(Lib, EntryPoint = "ddla_dsyrk_dev")]
    public static extern int ddla_dsyrk_dev(int uplo, int transa, long m, long k, in double alpha, IntPtr a, long lda, in double beta, IntPtr c, long ldc);

    [DllImport(Lib, EntryPoint = "ddla_dgemmt")]
    public static extern int ddla_dgemmt(int uploc, int transa, int transb, long m, long k, in double alpha, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb, in double beta, IntPtr c, long rsc, long csc);

    [DllImport(Lib, EntryPoint = "ddla_dgemmt_dev")]
    public static extern int ddla_dgemmt_dev(int uploc, int transa, int transb, long m, long k, in double alpha, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb, in double beta, IntPtr c, long rsc, long csc);

    [DllImport(Lib, EntryPoint = "ddla_dsyrk")]
    public static extern int ddla_dsyrk(int uploc, int transa, long m, long k, in double alpha, IntPtr a, long rsa, long csa, in double beta, IntPtr c, long rsc, long csc);

    [DllImport(Lib, EntryPoint = "ddla_dsyr2k")]
    public static extern int ddla_dsyr2k(int uploc, int transa, int transb, long m, long k, in double alpha, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb, in double beta, IntPtr c, long rsc, long csc);

    [DllImport(Lib, EntryPoint = "ddla_dsyr2k_dev")]
    public static extern int ddla_dsyr2k_dev(int uploc, int transa, int transb, long m, long k, in double alpha, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb, in double beta, IntPtr c, long rsc, long csc);

    [DllImport(Lib, EntryPoint = "ddla_dsymm")]
    public static extern int ddla_dsymm(int sidea, int uploa, int conja, int transb, long m, long n, in double alpha, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb, in double beta, IntPtr c, long rsc, long csc);

    [DllImport(Lib, EntryPoint = "ddla_dsymm_dev")]
    public static extern int ddla_dsymm_dev(int sidea, int uploa, int conja, int transb, long m, long n, in double alpha, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb, in double beta, IntPtr c, long rsc, long csc);

    [DllImport(Lib, EntryPoint = "ddla_dtrmm")]
    public static extern int ddla_dtrmm(int sidea, int uploa, int transa, int diaga, long m, long n, in double alpha, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb);

    [DllImport(Lib, EntryPoint = "ddla_dtrmm_dev")]
    public static extern int ddla_dtrmm_dev(int sidea, int uploa, int transa, int diaga, long m, long n, in double alpha, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb);

    [DllImport(Lib, EntryPoint = "ddla_dtrmm3")]
    public static extern int ddla_dtrmm3(int sidea, int uploa, int transa, int diaga, int transb, long m, long n, in double alpha, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb, in double beta, IntPtr c, long rsc, long csc);

    [DllImport(Lib, EntryPoint = "ddla_dtrmm3_dev")]
    public static extern int ddla_dtrmm3_dev(int sidea, int uploa, int transa, int diaga, int transb, long m, long n, in double alpha, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb, in double beta, IntPtr c, long rsc, long csc);

    [DllImport(Lib, EntryPoint = "ddla_dtrsm")]
    public static extern int ddla_dtrsm(int sidea, int uploa, int transa, int diaga, long m, long n, in double alpha, IntPtr a, long rsa, long csa, IntPtr b, long rsb, long csb);

    [DllImport(Lib, EntryPoint = "ddla_damaxv_dev")]
    public static extern int ddla_damaxv_dev(long n, IntPtr x, long incx, out long index);

    [DllImport(Lib, EntryPoint = "ddla_dgesv_mpir_dev")]
    public static extern int ddla_dgesv_mpir_dev(long n, long nrhs, IntPtr a, long lda, IntPtr b, long ldb, IntPtr x, long ldx, in MpirOpts opts, out MpirInfo info);

    [DllImport(Lib, EntryPoint = "ddla_dposv_mpir_dev")]
    public static extern int ddla_dposv_mpir_dev(long n, long nrhs, IntPtr a, long lda, IntPtr b, long ldb, IntPtr x, long ldx, in MpirOpts opts, out MpirInfo info);

    [DllImport(Lib, EntryPoint = "ddla_sgetrf_mp_dev")]
    public static extern int ddla_sgetrf_mp_dev(long n, IntPtr a, long lda, int prec, int pivoting, IntPtr perm);

    [DllImport(Lib, EntryPoint = "ddla_sgetf2_dev")]
    public static extern int ddla_sgetf2_dev(long m, long nb, IntPtr p, long ldp, IntPtr piv);

    [DllImport(Lib, EntryPoint = "ddla_dgetf2_dev")]
    public static extern int ddla_dgetf2_dev(long m, long nb, IntPtr p, long ldp, IntPtr piv);

    [DllImport(Lib, EntryPoint = "ddla_batch_create")]
    public static extern int ddla_batch_create(out IntPtr plan, int prec, long count, IntPtr n, IntPtr kind, IntPtr a, IntPtr piv, IntPtr info, int tune);

    [DllImport(Lib, EntryPoint = "ddla_batch_execute")]
    public static extern int ddla_batch_execute(IntPtr plan);

    [DllImport(Lib, EntryPoint = "ddla_batch_describe")]
    public static extern int ddla_batch_describe(IntPtr plan, StringBuilder buf, nuint len);

    [DllImport(Lib, EntryPoint = "ddla_batch_destroy")]
    public static extern int ddla_batch_destroy(IntPtr plan);

    /// <summary>The last error message of this thread's calls.</summary>
    public static string LastError()
    {
        var sb = new StringBuilder(1024);
        ddla_last_error(sb, (nuint)sb.Capacity);
        return sb.ToString();
    }
}
