// The level-3 part of DDLA's native Source class retargeted to libddla_hip.so: the same
// method names and argument lists as DDLA/BLAS/Native/3m.cs (bli_dgemm, bli_dgemmt, bli_dsymm,
// bli_dsyrk, bli_dsyr2k, bli_dtrmm, bli_dtrmm3, bli_dtrsm), so BlasProvider (DDLA/BLAS/3m.cs)
// and everything above it run on the GPU unchanged.  Host operands are copied to the device and
// back inside each call; strides follow BLIS (rs, cs in elements).
//
// Build with DDLA_HIP defined and the matching declarations of Native/3m.cs left out (wrap
// them in `#if !DDLA_HIP`).  The entry points return a status (0 or a DDLA_ERR_* code) where
// BLIS returns void; callers that discard it compile unchanged, and HipStatus.Check turns it
// into an exception.
#if DDLA_HIP
using System.Runtime.InteropServices;
using DDLA.Misc.Flags;
using ctype = double;

namespace DDLA.BLAS;

public static partial class Source
{
    public const string HipDLL = "ddla_hip";

    [DllImport(HipDLL, EntryPoint = "ddla_dgemm")]
    public static extern int GeMM(
        TransType transa, TransType transb,
        dim_t m, dim_t n, dim_t k,
        in ctype alpha,
        ref ctype a, inc_t rsa, inc_t csa,
        ref ctype b, inc_t rsb, inc_t csb,
        in ctype beta,
        ref ctype c, inc_t rsc, inc_t csc);

    [DllImport(HipDLL, EntryPoint = "ddla_dgemmt")]
    public static extern int GeMMT(
        UpLo uploc, TransType transa, TransType transb,
        dim_t m, dim_t k,
        in ctype alpha,
        ref ctype a, inc_t rsa, inc_t csa,
        ref ctype b, inc_t rsb, inc_t csb,
        in ctype beta,
        ref ctype c, inc_t rsc, inc_t csc);

    [DllImport(HipDLL, EntryPoint = "ddla_dsymm")]
    public static extern int SyMM(
        SideType sidea, UpLo uploa, ConjType conja, TransType transb,
        dim_t m, dim_t n,
        in ctype alpha,
        ref ctype a, inc_t rsa, inc_t csa,
        ref ctype b, inc_t rsb, inc_t csb,
        in ctype beta,
        ref ctype c, inc_t rsc, inc_t csc);

    [DllImport(HipDLL, EntryPoint = "ddla_dsyrk")]
    public static extern int SyRk(
        UpLo uploc, TransType transa,
        dim_t m, dim_t k,
        in ctype alpha,
        ref ctype a, inc_t rsa, inc_t csa,
        in ctype beta,
        ref ctype c, inc_t rsc, inc_t csc);

    [DllImport(HipDLL, EntryPoint = "ddla_dsyr2k")]
    public static extern int SyR2k(
        UpLo uploc, TransType transa, TransType transb,
        dim_t m, dim_t k,
        in ctype alpha,
        ref ctype a, inc_t rsa, inc_t csa,
        ref ctype b, inc_t rsb, inc_t csb,
        in ctype beta,
        ref ctype c, inc_t rsc, inc_t csc);

    [DllImport(HipDLL, EntryPoint = "ddla_dtrmm")]
    public static extern int TrMM(
        SideType sidea, UpLo uploa, TransType transa, DiagType diaga,
        dim_t m, dim_t n,
        in ctype alpha,
        ref ctype a, inc_t rsa, inc_t csa,
        ref ctype b, inc_t rsb, inc_t csb);

    [DllImport(HipDLL, EntryPoint = "ddla_dtrmm3")]
    public static extern int TrMM3(
        SideType sidea, UpLo uploa, TransType transa, DiagType diaga, TransType transb,
        dim_t m, dim_t n,
        in ctype alpha,
        ref ctype a, inc_t rsa, inc_t csa,
        ref ctype b, inc_t rsb, inc_t csb,
        in ctype beta,
        ref ctype c, inc_t rsc, inc_t csc);

    [DllImport(HipDLL, EntryPoint = "ddla_dtrsm")]
    public static extern int TrSM(
        SideType sidea, UpLo uploa, TransType transa, DiagType diaga,
        dim_t m, dim_t n,
        in ctype alpha,
        ref ctype a, inc_t rsa, inc_t csa,
        ref ctype b, inc_t rsb, inc_t csb);

    /// <summary>index = IndexOfMax(Abs(x)) on the GPU (bli_damaxv, Native/1v.cs:26-31): NaN is
    /// smaller than any value, the first of equal maxima wins.</summary>
    public static void AmaxDevice(dim_t n, System.IntPtr x, inc_t incx, out dim_t index)
        => DDLA.Hip.HipStatus.Check(DDLA.Hip.HipNative.ddla_damaxv_dev(n, x, incx, out index));
}
#endif
