/*
 * ddla_native.h -- C ABI of the MI355X (gfx950) dense linear-algebra backend for DDLA.
 *
 * The library (ddla_amd/libddla_hip.so) replaces, for DDLA's dense hot path, the BLIS
 * entry points that DDLA reaches through P/Invoke (`static partial class Source`,
 * DDLA/BLAS/Native/utils.cs:8-17) and the managed blocked factorizations built on them
 * (DDLA/Factorizations/LU.cs, DDLA/Factorizations/Cholesky.cs).
 *
 * Conventions (all taken from the P/Invoke declarations the entry points replace):
 *   - dim_t / inc_t are 64-bit signed integers (DDLA/Alias.cs:1-2).
 *   - A matrix argument is (pointer to element (0,0), row stride rs, column stride cs) in
 *     elements, like MatrixView (DDLA/Core/MatrixView.cs:22-31,483-485).  Host-pointer
 *     entry points accept any strides; device-pointer (`_dev`) entry points take row-major
 *     device matrices with a leading dimension ld (element (i,j) at p[i*ld + j]).
 *   - Flag values are DDLA's BLIS-valued enums (the files in DDLA/Misc/Flags).
 *   - Scalars are passed by pointer where the replaced BLIS call does (`in double`).
 *   - Every entry point returns int: 0 success, > 0 LAPACK-style info (1-based index of
 *     the first zero pivot / non-positive Cholesky pivot), < 0 an error code below; the
 *     message of the last error is available from ddla_last_error().
 *   - LU pivots use DDLA's relative encoding: piv[i] = p - i where row p was swapped with
 *     row i at step i, 0 = no swap (LU.cs:367-374,386-391).
 *   - Device entry points are ordered on the library stream (ddla_set_stream) and return
 *     after the work is queued, except those that return an info value or fill a result
 *     struct, which synchronise the stream before returning.  Host-pointer entry points
 *     are synchronous and keep no host pointer after they return.
 *   - One library context per process; calls from several host threads must be
 *     serialised by the caller.
 */
#ifndef DDLA_NATIVE_H
#define DDLA_NATIVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t ddla_dim_t; /* dim_t, DDLA/Alias.cs:1 */
typedef int64_t ddla_inc_t; /* inc_t, DDLA/Alias.cs:2 */

/* TransType (DDLA/Misc/Flags/TransType.cs:3-9) */
#define DDLA_NO_TRANSPOSE 0x00
#define DDLA_TRANSPOSE 0x08
#define DDLA_CONJ_NO_TRANSPOSE 0x10
#define DDLA_CONJ_TRANSPOSE 0x18
/* UpLo (DDLA/Misc/Flags/UpLo.cs:3-9) */
#define DDLA_ZEROS 0x00
#define DDLA_LOWER 0xC0
#define DDLA_UPPER 0x60
#define DDLA_DENSE 0xE0
/* DiagType (DDLA/Misc/Flags/DiagType.cs:3-7) */
#define DDLA_NONUNIT_DIAG 0x000
#define DDLA_UNIT_DIAG 0x100
/* SideType (DDLA/Misc/Flags/SideType.cs:3-7) */
#define DDLA_LEFT 0
#define DDLA_RIGHT 1

/* status codes */
#define DDLA_SUCCESS 0
#define DDLA_ERR_INVALID_ARG (-1)
#define DDLA_ERR_HIP (-2)
#define DDLA_ERR_NO_DEVICE (-3)
#define DDLA_ERR_ALLOC (-4)
#define DDLA_ERR_UNSUPPORTED (-5)
#define DDLA_ERR_NOT_CONVERGED (-6)

/* factor precisions of the mixed-precision solvers */
#define DDLA_PREC_FP64 0
#define DDLA_PREC_FP32 1
#define DDLA_PREC_FP16 2
#define DDLA_PREC_BF16 3

/* ---------------------------------------------------------------- runtime ---------- */
/* No counterpart in the reference (BLIS needs no context); explicit set-up so a P/Invoke
 * caller controls device selection and tear-down. */
int ddla_init(int device);
int ddla_finalize(void);
/* Copies the message of the last error (NUL-terminated, truncated to len). */
int ddla_last_error(char* buf, size_t len);
/* Library stream: a hipStream_t cast to void*; NULL selects the library's own stream. */
int ddla_set_stream(void* stream);
int ddla_get_stream(void** stream);
int ddla_sync(void);
/* Library version as 10000*major + 100*minor + patch. */
int ddla_version(void);
/* No counterpart in the reference.  The LU panel and the persistent triangular solve run as
 * workgroups that wait on each other's results; each wait gives up after `rounds` poll rounds
 * (0 restores the default, 2^24, many seconds), the launch's workgroups leave, and the call
 * returns DDLA_ERR_HIP instead of hanging -- e.g. when other work keeps the workgroups from
 * becoming co-resident. */
int ddla_set_wait_limit(uint64_t rounds);

/* Per-kernel timing, no counterpart in the reference.  While enabled, every launch of the
 * library's kernels is bracketed by HIP events on its stream; the totals per kernel name
 * (launch count, GPU milliseconds, algorithmic flops and bytes) are read back with
 * ddla_profile_read.  idx < 0 returns the number of entries. */
int ddla_profile_enable(int on);
int ddla_profile_reset(void);
int ddla_profile_read(int idx, char* name, size_t len, int64_t* count, double* total_ms,
                      double* flops, double* bytes);
/* Restricts the timing to the launches of one kernel name (null or "": every kernel), so that
 * a timed region can carry the events of the kernel it is judged by and nothing else. */
int ddla_profile_only(const char* name);

/* ------------------------------------------------------- device buffers ---------- */
/* Device-resident matrices (the boundary's replacement for pinned managed arrays). */
int ddla_malloc(void** dptr, size_t bytes);
int ddla_free(void* dptr);
int ddla_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes);
int ddla_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes);
int ddla_memcpy_d2d(void* dst_dev, const void* src_dev, size_t bytes);
/* Strided host view (m x n, rs, cs) <-> row-major device matrix with leading dim ld. */
int ddla_dmat_set(ddla_dim_t m, ddla_dim_t n, const double* host, ddla_inc_t rs, ddla_inc_t cs,
                  double* dev, ddla_inc_t ld);
int ddla_dmat_get(ddla_dim_t m, ddla_dim_t n, const double* dev, ddla_inc_t ld, double* host,
                  ddla_inc_t rs, ddla_inc_t cs);
/* Fills an m x n row-major device matrix with the SplitMix64 stream mapped to [-1, 1):
 * element k = i*n + j gets (splitmix64(seed + (k+1)*0x9E3779B97F4A7C15) >> 11)*2^-52 - 1.
 * Replaces Rand/randm (DDLA/BLAS/Managed/utils.cs:145-233 -> bli_drandm) for device data. */
int ddla_drandm_dev(ddla_dim_t m, ddla_dim_t n, double* dev, ddla_inc_t ld, uint64_t seed);

/* ------------------------------------------------------------------- GEMM ---------- */
/* C = alpha*op(A)*op(B) + beta*C with BLIS-style strides; host pointers; synchronous.
 * Same argument list as bli_dgemm (DDLA/BLAS/Native/3m.cs:12-24), the call behind
 * BlasProvider.GeMM (DDLA/BLAS/3m.cs:10-42).  beta == 0 overwrites C without reading it
 * (Managed/3m.cs:32 -> Managed/1m.cs:265-269); k == 0 only scales C. */
int ddla_dgemm(int transa, int transb, ddla_dim_t m, ddla_dim_t n, ddla_dim_t k,
               const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa,
               const double* b, ddla_inc_t rsb, ddla_inc_t csb, const double* beta, double* c,
               ddla_inc_t rsc, ddla_inc_t csc);
/* Same operation on device pointers (any strides; unit stride along rows or columns of
 * every operand takes the fast path); stream-ordered. */
int ddla_dgemm_dev(int transa, int transb, ddla_dim_t m, ddla_dim_t n, ddla_dim_t k,
                   const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa,
                   const double* b, ddla_inc_t rsb, ddla_inc_t csb, const double* beta,
                   double* c, ddla_inc_t rsc, ddla_inc_t csc);

/* ------------------------------------------------------ LU (partial pivoting) ------ */
/* Blocked right-looking PLU of a row-major n x n device matrix, block 128, in place.
 * Replaces LU.PLUDecompose (DDLA/Factorizations/LU.cs:251-264, PLUDecBlock :289-342,
 * PLUDecUnblock :344-374).  dpiv: device int32[n], relative pivots.  Returns info
 * (first zero pivot, 1-based; the reference throws from InvScal, DDLA/BLAS/1v.cs:99). */
int ddla_dgetrf_dev(ddla_dim_t n, double* a, ddla_inc_t lda, int32_t* dpiv);
/* ddla_dgetrf_dev without any device allocation.  No counterpart in the reference: for
 * callers that cannot call hipMalloc (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  Same kernels, pivots and factors as ddla_dgetrf_dev;
 * the scratch (panel exchange area, info word, row lists: about 1.6 MiB) lives in a static
 * device array that each GPU holds until the process ends, set up by ddla_init.  Orders above
 * 32768 return DDLA_ERR_INVALID_ARG.  All calls on one GPU share that array: calls that
 * overlap on one GPU race. */
int ddla_dgetrf_noalloc_dev(ddla_dim_t n, double* a, ddla_inc_t lda, int32_t* dpiv);
/* Solves A X = B (trans == DDLA_NO_TRANSPOSE) or A^T X = B with the factors of
 * ddla_dgetrf_dev; B (n x nrhs row-major, device) is overwritten with X.
 * Replaces LU.PLUSolve (LU.cs:266-287, PLUSolveInternel :431-465).  The pivots are checked on
 * the device first (a pivot that would swap outside the matrix returns DDLA_ERR_INVALID_ARG,
 * where the reference's SwapRow throws), which synchronises the stream; with at most 8
 * right-hand sides the stream is synchronised again at the end to report a triangular-solve
 * wait that gave up (ddla_set_wait_limit). */
int ddla_dgetrs_dev(int trans, ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t lda,
                    const int32_t* dpiv, double* b, ddla_inc_t ldb);
/* Host-pointer forms with strided views (synchronous). */
int ddla_dgetrf(ddla_dim_t n, double* a, ddla_inc_t rsa, ddla_inc_t csa, int32_t* piv);
int ddla_dgetrs(int trans, ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t rsa,
                ddla_inc_t csa, const int32_t* piv, double* b, ddla_inc_t rsb, ddla_inc_t csb);
/* Row interchanges with relative pivots: side DDLA_LEFT swaps rows, DDLA_RIGHT columns;
 * forward applies P, backward P^-1.  Replaces LU.ApplyPiv (LU.cs:237-249, :376-429).  Pivots
 * are range-checked on the device first, as in ddla_dgetrs_dev. */
int ddla_dlaswp_dev(int side, int forward, ddla_dim_t m, ddla_dim_t n, double* x, ddla_inc_t ldx,
                    const int32_t* dpiv, ddla_dim_t npiv);

/* LU without pivoting (block 128, blocked iff n > 128), in place.  Replaces
 * LU.LUDecompose (DDLA/Factorizations/LU.cs:133-142, LUDecBlock :169-186, LUDecUnblock
 * :188-203).  Returns info (first zero pivot, 1-based). */
int ddla_dgetrf_nopiv_dev(ddla_dim_t n, double* a, ddla_inc_t lda);
/* Solves A X = B or A^T X = B with the factors of ddla_dgetrf_nopiv_dev.  Replaces
 * LU.LUSolve (LU.cs:144-167, LUSolveInternal :205-231). */
int ddla_dgetrs_nopiv_dev(int trans, ddla_dim_t n, ddla_dim_t nrhs, const double* a,
                          ddla_inc_t lda, double* b, ddla_inc_t ldb);

/* ----------------------------------------------------------------- Cholesky -------- */
/* Blocked Cholesky (block 128) of a row-major n x n SPD device matrix in place; only the
 * uplo triangle is read and written.  Replaces Cholesky.CholDecompose
 * (DDLA/Factorizations/Cholesky.cs:52-69, lower :127-160, upper :162-208).  Returns info
 * (1-based index of the first pivot that is not > 0; Cholesky.cs:228-236). */
int ddla_dpotrf_dev(int uplo, ddla_dim_t n, double* a, ddla_inc_t lda);
/* Solves A X = B with the factor of ddla_dpotrf_dev (Cholesky.cs:102-123, :210-226). */
int ddla_dpotrs_dev(int uplo, ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t lda,
                    double* b, ddla_inc_t ldb);
int ddla_dpotrf(int uplo, ddla_dim_t n, double* a, ddla_inc_t rsa, ddla_inc_t csa);
int ddla_dpotrs(int uplo, ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t rsa,
                ddla_inc_t csa, double* b, ddla_inc_t rsb, ddla_inc_t csb);

/* ---------------------------------------------------------------------- LDL^T -------- */
/* Blocked LDL^T without pivoting (block 128) of a row-major n x n symmetric device matrix in
 * place: lower A = L D L^T, upper A = U^T D U, the unit triangle stored below (above) the
 * diagonal and D on it; only the uplo triangle is read and written.  Replaces
 * LDLT.LDLTDecompose (DDLA/Factorizations/LDLT.cs:56-73; lower :141-202, upper :204-265).
 * Returns info = 1-based index of the first zero pivot d_k (the reference throws
 * LinalgException for a zero pivot, LDLT.cs:148-152, but does not test the last pivot of
 * each 128-block; every pivot is tested here). */
int ddla_dldlt_dev(int uplo, ddla_dim_t n, double* a, ddla_inc_t lda);
/* Solves A X = B with the factor of ddla_dldlt_dev (LDLT.LDLTSolve, LDLT.cs:105-126,
 * :267-295). */
int ddla_dldlt_solve_dev(int uplo, ddla_dim_t n, ddla_dim_t nrhs, const double* a,
                         ddla_inc_t lda, double* b, ddla_inc_t ldb);
/* Host forms with (row stride, column stride) views. */
int ddla_dldlt(int uplo, ddla_dim_t n, double* a, ddla_inc_t rsa, ddla_inc_t csa);
int ddla_dldlt_solve(int uplo, ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t rsa,
                     ddla_inc_t csa, double* b, ddla_inc_t rsb, ddla_inc_t csb);

/* ------------------------------------------------------------------------- QR ------- */
/* Blocked Householder QR in the UT form (block 256).  T holds ddla_qr_t_rows(m, n) =
 * min(256, max(m, n)) rows and n columns: block k (columns k .. k+b-1) keeps its b x b upper
 * triangular factor in T[0:b, k:k+b] (tau_j on the diagonal, u_i^T u_j above); other entries of
 * T are unspecified.  Replaces QR.CreateT / QR.QRDecompose (DDLA/Factorizations/QR.cs:124-149,
 * :326-395): A overwritten with R on and above the diagonal and the Householder vectors u2
 * below (implicit 1 on the diagonal; H = I - u u^T / tau, HouseHolder.cs:34-56).  The column
 * norm is sqrt of the sum of squares (the reference uses BLIS's scaled norm); m <= 32768. */
int ddla_qr_t_rows(ddla_dim_t m, ddla_dim_t n);
int ddla_dgeqrf_dev(ddla_dim_t m, ddla_dim_t n, double* a, ddla_inc_t lda, double* t,
                    ddla_inc_t ldt);
/* B (m x nrhs) <- Q^T B (trans = DDLA_TRANSPOSE; QR.ApplyQlhfc, QR.cs:423-464) or Q B
 * (DDLA_NO_TRANSPOSE; QR.ApplyQlnfc, :466-521), with the factor of ddla_dgeqrf_dev. */
int ddla_dqr_apply_dev(int trans, ddla_dim_t m, ddla_dim_t n, const double* a, ddla_inc_t lda,
                       const double* t, ddla_inc_t ldt, ddla_dim_t nrhs, double* b,
                       ddla_inc_t ldb);
/* Least squares X (n x nrhs) = argmin |A X - B| for m >= n: X = R^-1 (Q^T B)[0:n]; B is left
 * unchanged (QR.Solve, QR.cs:151-167, :397-409). */
int ddla_dqr_solve_dev(ddla_dim_t m, ddla_dim_t n, ddla_dim_t nrhs, const double* a,
                       ddla_inc_t lda, const double* t, ddla_inc_t ldt, const double* b,
                       ddla_inc_t ldb, double* x, ddla_inc_t ldx);
/* Host forms with (row stride, column stride) views. */
int ddla_dgeqrf(ddla_dim_t m, ddla_dim_t n, double* a, ddla_inc_t rsa, ddla_inc_t csa,
                double* t, ddla_inc_t rst, ddla_inc_t cst);
int ddla_dqr_solve(ddla_dim_t m, ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t rsa,
                   ddla_inc_t csa, const double* t, ddla_inc_t rst, ddla_inc_t cst,
                   const double* b, ddla_inc_t rsb, ddla_inc_t csb, double* x, ddla_inc_t rsx,
                   ddla_inc_t csx);

/* ------------------------------------------------------------- tensor contraction ---- */
/* C = beta C + alpha contract(expr, A, B) for strided FP64 device tensors (lengths and
 * element strides per dimension, at most 16 dimensions).  expr is "<A symbols>,<B
 * symbols>-><C symbols>", one character per dimension.  Replaces MArray.Contract(beta,
 * result, expression, alpha, left, right) (DDLA/Einsum/MArray.Contract.cs:35-125, the
 * block-scatter contraction :252-405) with the index rules of IndiceUtils (Utilities/
 * IndiceUtils.cs:18-130): a repeated symbol in one tensor is its diagonal; symbols of A and B
 * only are summed products, of A alone or B alone summed, of C alone broadcast.  beta scales
 * every element of C first (beta = 0 stores 0); alpha = 0 leaves C scaled only.  Errors
 * (DDLA_ERR_INVALID_ARG) for a malformed expression, a rank that does not match its symbols,
 * or a symbol with different lengths. */
int ddla_dcontract_dev(const char* expr, double alpha, const double* a, int na,
                       const int64_t* len_a, const int64_t* str_a, const double* b, int nb,
                       const int64_t* len_b, const int64_t* str_b, double beta, double* c,
                       int nc, const int64_t* len_c, const int64_t* str_c);

/* ------------------------------------------------------------ level-3 pieces ------- */
/* B = alpha * op(A)^-1 B (left) or B op(A)^-1 (right), row-major device matrices with leading
 * dimensions: the device-resident form of bli_dtrsm (DDLA/BLAS/Native/3m.cs:151-162) for the
 * factorizations' row-major buffers; ddla_dtrsm below has bli_dtrsm's own argument list. */
int ddla_dtrsm_dev(int side, int uplo, int transa, int diag, ddla_dim_t m, ddla_dim_t n,
                   const double* alpha, const double* a, ddla_inc_t lda, double* b,
                   ddla_inc_t ldb);
/* uplo triangle of C = beta*C + alpha*op(A) op(A)^T, row-major device matrices with leading
 * dimensions: the device-resident form of bli_dsyrk (DDLA/BLAS/Native/3m.cs:96-106);
 * ddla_dsyrk below has bli_dsyrk's own argument list. */
int ddla_dsyrk_dev(int uplo, int transa, ddla_dim_t m, ddla_dim_t k, const double* alpha,
                   const double* a, ddla_inc_t lda, const double* beta, double* c,
                   ddla_inc_t ldc);
/* Level-3 BLAS with exactly BLIS's argument lists (strided views, scalars by pointer), so the
 * C# `Source` declarations of DDLA/BLAS/Native/3m.cs retarget by changing only the library
 * name and entry point.  Host-pointer forms are synchronous; `_dev` forms take device
 * pointers with any strides and are ordered on the library stream.  conja is ConjType
 * (DDLA/Misc/Flags/ConjType.cs: 0 or 0x10), the identity for real data.
 * C (m x m, uploc triangle only) = beta*C + alpha*op(A)*op(B), op(A) m x k, op(B) k x m.
 * Replaces bli_dgemmt (Native/3m.cs:26-38) behind BlasProvider.GeMMT (BLAS/3m.cs:44-70). */
int ddla_dgemmt(int uploc, int transa, int transb, ddla_dim_t m, ddla_dim_t k, const double* alpha,
                const double* a, ddla_inc_t rsa, ddla_inc_t csa, const double* b, ddla_inc_t rsb,
                ddla_inc_t csb, const double* beta, double* c, ddla_inc_t rsc, ddla_inc_t csc);
int ddla_dgemmt_dev(int uploc, int transa, int transb, ddla_dim_t m, ddla_dim_t k,
                    const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa,
                    const double* b, ddla_inc_t rsb, ddla_inc_t csb, const double* beta,
                    double* c, ddla_inc_t rsc, ddla_inc_t csc);
/* uploc triangle of C = beta*C + alpha*op(A)*op(A)^T.  Replaces bli_dsyrk (Native/3m.cs:96-106)
 * behind BlasProvider.SyRk (BLAS/3m.cs:118-138). */
int ddla_dsyrk(int uploc, int transa, ddla_dim_t m, ddla_dim_t k, const double* alpha,
               const double* a, ddla_inc_t rsa, ddla_inc_t csa, const double* beta, double* c,
               ddla_inc_t rsc, ddla_inc_t csc);
/* uploc triangle of C = beta*C + alpha*op(A)*op(B)^T + alpha*op(B)*op(A)^T, op(A), op(B) m x k.
 * Replaces bli_dsyr2k (Native/3m.cs:108-120) behind BlasProvider.SyR2k (BLAS/3m.cs:149-172). */
int ddla_dsyr2k(int uploc, int transa, int transb, ddla_dim_t m, ddla_dim_t k, const double* alpha,
                const double* a, ddla_inc_t rsa, ddla_inc_t csa, const double* b, ddla_inc_t rsb,
                ddla_inc_t csb, const double* beta, double* c, ddla_inc_t rsc, ddla_inc_t csc);
int ddla_dsyr2k_dev(int uploc, int transa, int transb, ddla_dim_t m, ddla_dim_t k,
                    const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa,
                    const double* b, ddla_inc_t rsb, ddla_inc_t csb, const double* beta,
                    double* c, ddla_inc_t rsc, ddla_inc_t csc);
/* C (m x n) = beta*C + alpha*S*op(B) (sidea LEFT, S m x m) or alpha*op(B)*S (RIGHT, S n x n),
 * S symmetric with only its uploa triangle read.  Replaces bli_dsymm (Native/3m.cs:81-94)
 * behind BlasProvider.SyMM (BLAS/3m.cs:83-106). */
int ddla_dsymm(int sidea, int uploa, int conja, int transb, ddla_dim_t m, ddla_dim_t n,
               const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa,
               const double* b, ddla_inc_t rsb, ddla_inc_t csb, const double* beta, double* c,
               ddla_inc_t rsc, ddla_inc_t csc);
int ddla_dsymm_dev(int sidea, int uploa, int conja, int transb, ddla_dim_t m, ddla_dim_t n,
                   const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa,
                   const double* b, ddla_inc_t rsb, ddla_inc_t csb, const double* beta, double* c,
                   ddla_inc_t rsc, ddla_inc_t csc);
/* B (m x n) = alpha*op(T)*B (LEFT) or alpha*B*op(T) (RIGHT), T the uploa triangle of A with
 * a unit diagonal when diaga is DDLA_UNIT_DIAG.  Replaces bli_dtrmm (Native/3m.cs:122-133)
 * behind BlasProvider.TrMM (BLAS/3m.cs:185-204); op(T) honours transa as BLIS does (the
 * managed TrMM, BLAS/Managed/3m.cs:259-287, does not apply aTrans). */
int ddla_dtrmm(int sidea, int uploa, int transa, int diaga, ddla_dim_t m, ddla_dim_t n,
               const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa, double* b,
               ddla_inc_t rsb, ddla_inc_t csb);
int ddla_dtrmm_dev(int sidea, int uploa, int transa, int diaga, ddla_dim_t m, ddla_dim_t n,
                   const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa, double* b,
                   ddla_inc_t rsb, ddla_inc_t csb);
/* C (m x n) = beta*C + alpha*op(T)*op(B) (LEFT) or alpha*op(B)*op(T) (RIGHT).  Replaces
 * bli_dtrmm3 (Native/3m.cs:135-149) behind BlasProvider.TrMM3 (BLAS/3m.cs:225-250). */
int ddla_dtrmm3(int sidea, int uploa, int transa, int diaga, int transb, ddla_dim_t m,
                ddla_dim_t n, const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa,
                const double* b, ddla_inc_t rsb, ddla_inc_t csb, const double* beta, double* c,
                ddla_inc_t rsc, ddla_inc_t csc);
int ddla_dtrmm3_dev(int sidea, int uploa, int transa, int diaga, int transb, ddla_dim_t m,
                    ddla_dim_t n, const double* alpha, const double* a, ddla_inc_t rsa,
                    ddla_inc_t csa, const double* b, ddla_inc_t rsb, ddla_inc_t csb,
                    const double* beta, double* c, ddla_inc_t rsc, ddla_inc_t csc);
/* Solves op(T)*X = alpha*B (LEFT) or X*op(T) = alpha*B (RIGHT), X over B, strided host views.
 * Replaces bli_dtrsm (Native/3m.cs:151-162) behind BlasProvider.TrSM (BLAS/3m.cs:287-305). */
int ddla_dtrsm(int sidea, int uploa, int transa, int diaga, ddla_dim_t m, ddla_dim_t n,
               const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa, double* b,
               ddla_inc_t rsb, ddla_inc_t csb);
/* Index of the first max |x_i| (NaN smallest; empty -> 0), written to *index (host).
 * Replaces bli_damaxv (DDLA/BLAS/Native/1v.cs:26-31) behind BlasProvider.AMax
 * (DDLA/BLAS/1v.cs:30-41). */
int ddla_damaxv_dev(ddla_dim_t n, const double* x, ddla_inc_t incx, ddla_dim_t* index);

/* ------------------------------------------------ mixed precision + refinement ----- */
/* No counterpart in the reference: factor in a lower precision on the matrix cores, then
 * recover FP64 accuracy by iterative refinement on the GPU (residual, correction and
 * convergence test in FP64). */
#define DDLA_PIVOT_PARTIAL 0
#define DDLA_PIVOT_BUTTERFLY 1
#define DDLA_PIVOT_TOURNAMENT 2

typedef struct ddla_mpir_opts {
  int32_t factor_prec;   /* DDLA_PREC_FP16 / _BF16 / _FP32 / _FP64 */
  int32_t max_iters;     /* refinement steps before giving up (default 50 if <= 0) */
  double tol;            /* target normwise backward error (default 1e-14 if <= 0) */
  int32_t gmres;         /* must be 0 (classic refinement); other values are rejected with
                            DDLA_ERR_UNSUPPORTED */
  int32_t terms;         /* low-precision parts per FP32 operand in the trailing updates:
                            1 = one FP16/BF16 rounding, 2 or 3 = split x = h + m (+ l) with
                            the cross products summed; 0 = default (FP16: 2, BF16: 3) */
  int32_t pivoting;      /* LU only (DDLA_PIVOT_*): PARTIAL = partial pivoting (the FP64
                            path's pivot rule); BUTTERFLY = two-sided random butterfly
                            transform, then LU without pivoting (no global pivot search);
                            TOURNAMENT = tournament pivoting in each 128-column panel (pivot
                            rows chosen by a reduction tree of small partial-pivoting
                            eliminations, as in CALU).  BUTTERFLY and TOURNAMENT fall back to
                            partial pivoting when their factor breaks down or does not refine
                            (fallback = 2 / 3) */
  int32_t reserved;      /* must be 0 */
  const double* x_ref;   /* optional (device, n x nrhs, leading dimension ld_xref): a reference
                            solution; info->forward_error is measured against it */
  int64_t ld_xref;
} ddla_mpir_opts;

typedef struct ddla_mpir_info {
  int32_t status;        /* 0 converged, > 0 factor info, < 0 error / not converged */
  int32_t iters;         /* refinement steps (GMRES inner steps counted in inner_iters) */
  int32_t inner_iters;
  int32_t fallback;      /* 1: the low-precision Cholesky factor was not positive definite
                            and an FP32 factor was used; 2: the butterfly LU fell back to
                            partial pivoting; 3: the tournament-pivoting LU did */
  double backward_error; /* max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf) */
  double factor_ms;      /* GPU time of the factorization */
  double refine_ms;      /* GPU time of the refinement */
  double forward_error;  /* max over columns of ||x - x_ref||_inf / ||x_ref||_inf when opts->x_ref
                            is given, else -1 */
} ddla_mpir_info;

/* Solves A X = B for a row-major n x n FP64 device matrix A (not modified) and n x nrhs
 * B, writing X (n x nrhs, ldx).  LU with partial pivoting in the factor precision.  The
 * residual, the correction and the convergence test run in FP64 on the GPU; the host queues
 * refinement steps in groups and reads the test's state once per group (steps queued after
 * convergence do nothing). */
int ddla_dgesv_mpir_dev(ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t lda,
                        const double* b, ddla_inc_t ldb, double* x, ddla_inc_t ldx,
                        const ddla_mpir_opts* opts, ddla_mpir_info* info);
/* Same for SPD A with a lower Cholesky factor; falls back to an FP32 factor when the
 * low-precision factor is not positive definite. */
int ddla_dposv_mpir_dev(ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t lda,
                        const double* b, ddla_inc_t ldb, double* x, ddla_inc_t ldx,
                        const ddla_mpir_opts* opts, ddla_mpir_info* info);
/* The low-precision factor of ddla_dgesv_mpir_dev on its own: LU of a row-major n x n FP32
 * device matrix in place (unit L below the diagonal, U on and above), trailing updates in
 * prec (DDLA_PREC_FP32 / _FP16 / _BF16), pivoting DDLA_PIVOT_PARTIAL or _TOURNAMENT.
 * perm (device, n entries): row i of the factored matrix is row perm[i] of the input.
 * Returns > 0 (first zero pivot column + 1) for a singular factor.  No reference
 * counterpart (a preconditioner for callers running their own refinement). */
int ddla_sgetrf_mp_dev(ddla_dim_t n, float* a, ddla_inc_t lda, int32_t prec, int32_t pivoting,
                       int32_t* perm);
/* The panel of ddla_sgetrf_mp_dev with partial pivoting on its own: PLUDecUnblock
 * (DDLA/Factorizations/LU.cs:344-374, GetAndApplyPivUnblock's AMax rule) in FP32 on a
 * row-major m x nb device panel in place, 1 <= nb <= min(m, 128), m <= 32768; relative
 * pivots (nb entries, device) as ddla_dgetrf_dev.  Returns > 0 (first zero pivot column + 1)
 * where the reference's panel stops.  The reference has no FP32 path; the panel is exposed for
 * callers that build their own blocked FP32 LU and for parity tests. */
int ddla_sgetf2_dev(ddla_dim_t m, ddla_dim_t nb, float* p, ddla_inc_t ldp, int32_t* piv);
/* PLUDecUnblock (DDLA/Factorizations/LU.cs:344-374) on its own: the FP64 LU panel of
 * ddla_dgetrf_dev on a row-major m x nb device panel in place, 1 <= nb <= min(m, 128),
 * m <= 32768; relative pivots (nb entries, device).  Returns > 0 (first zero pivot column + 1)
 * where the reference's panel stops. */
int ddla_dgetf2_dev(ddla_dim_t m, ddla_dim_t nb, double* p, ddla_inc_t ldp, int32_t* piv);

/* ------------------------------------------------ batched uneven factorizations ----- */
/* No counterpart in the reference (DDLA factors one matrix per call: LU.PLUDecompose,
 * DDLA/Factorizations/LU.cs:251-264, Cholesky.CholDecompose, Cholesky.cs:52-69).  Many
 * independent square problems of order 1..256, FP64 or FP32, packed row-major one after
 * another in one device buffer: problem i (order n[i], leading dimension n[i]) starts at
 * element sum_{j<i} n[j]^2, its pivots (LU only, relative encoding as ddla_dgetrf_dev) at
 * sum_{j<i} n[j] in piv, its status in info[i] (0, or the 1-based column of the first zero
 * pivot / non-positive Cholesky pivot, where that problem stopped).
 *
 * A plan binds the buffers, sorts and balances the problems over one persistent launch,
 * chooses the panel width and workgroups per CU, and captures the launch in a hipGraph.
 * tune = 0: a fixed default shape.  tune = 1: the shape stored for this batch's bucket (device
 * architecture, precision, largest order rounded up to a power of two, log2 of the count,
 * LU / Cholesky / mixed) in the tuning cache, else every candidate is timed on a copy of the
 * batch (the input is not modified) and the winner is stored; tune = 2: always time, and store.
 * The cache is a JSON file at $DDLA_TUNE_CACHE, else $XDG_CACHE_HOME/ddla_amd/tune.json, else
 * $HOME/.cache/ddla_amd/tune.json; an unreadable or unwritable file only disables it.
 * ddla_batch_execute replays the graph on the library stream (asynchronous; ddla_sync waits).
 * n and kind are host arrays; a, piv and info are device buffers. */
#define DDLA_BATCH_LU 0   /* LU with partial pivoting */
#define DDLA_BATCH_CHOL 1 /* lower Cholesky (the upper triangle is not touched) */
typedef struct ddla_batch_s* ddla_batch_t;
int ddla_batch_create(ddla_batch_t* plan, int prec, int64_t count, const int32_t* n,
                      const int32_t* kind, void* a, int32_t* piv, int32_t* info, int tune);
int ddla_batch_execute(ddla_batch_t plan);
/* JSON text: chosen panel width, workgroups per CU, workgroups, flop and byte counts, the
 * autotuner's candidate timings (ms), the bucket key and whether the shape came from the cache */
int ddla_batch_describe(ddla_batch_t plan, char* buf, size_t len);
int ddla_batch_destroy(ddla_batch_t plan);

#ifdef __cplusplus
}
#endif

#endif /* DDLA_NATIVE_H */
