// Level-3 BLAS with BLIS argument lists: GEMMT, SYRK, SYR2K, SYMM, TRMM, TRMM3, TRSM.
//
// These replace the P/Invoke declarations of DDLA/BLAS/Native/3m.cs (bli_dgemmt :26-38,
// bli_dsymm :81-94, bli_dsyrk :96-106, bli_dsyr2k :108-120, bli_dtrmm :122-133, bli_dtrmm3
// :135-149, bli_dtrsm :151-162), reached through the facade BlasProvider (DDLA/BLAS/3m.cs).
// Every operation runs on the MFMA GEMM of gemm_mfma.hip:
//   * GEMMT / SYRK / SYR2K: triangle-masked GEMM tiles (only tiles that meet the stored
//     triangle run, and elements outside it are neither read nor written);
//   * SYMM / TRMM / TRMM3: the symmetric or triangular operand is first expanded into a dense
//     row-major copy (O(k^2) bytes against the product's O(k^2 n) flops), then one GEMM;
//   * TRSM: the blocked triangular solve of trsm.hip.
// Host-pointer forms upload each strided view into a contiguous row-major device buffer, run
// the device form and copy the result back (synchronous, no host pointer kept).  Device forms
// take device pointers with any strides and are ordered on the library stream.
#include <algorithm>

#include "ddla_internal.hpp"

using namespace ddla;

namespace {

constexpr int CONJ = 0x10;  // ConjType.Conj (DDLA/Misc/Flags/ConjType.cs); identity for real data

bool valid_uplo(int u) { return u == DDLA_LOWER || u == DDLA_UPPER; }
bool valid_side(int s) { return s == DDLA_LEFT || s == DDLA_RIGHT; }
bool valid_diag(int d) { return d == DDLA_UNIT_DIAG || d == DDLA_NONUNIT_DIAG; }
int tri_of(int uplo) { return uplo == DDLA_LOWER ? 1 : 2; }

// out (k x k, row-major) from the stored triangle of a (strides rs, cs):
//   SYM:  the symmetric matrix whose `lower` (or upper) triangle is stored;
//   TRI:  the triangle itself, zeros outside it, ones on the diagonal when `unit`.
template <bool SYM>
__global__ void expand_kernel(int64_t k, const double* __restrict__ a, int64_t rs, int64_t cs,
                              bool lower, bool unit, double* __restrict__ out) {
  const int64_t total = k * k;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / k, j = t % k;
    const bool stored = lower ? j <= i : j >= i;
    double v;
    if (SYM) {
      v = stored ? a[i * rs + j * cs] : a[j * rs + i * cs];
    } else {
      v = stored ? a[i * rs + j * cs] : 0.0;
      if (unit && i == j) v = 1.0;
    }
    out[t] = v;
  }
}

int expand(bool sym, int64_t k, const double* a, int64_t rs, int64_t cs, bool lower, bool unit,
           double** out) {
  void* ws = nullptr;
  DDLA_RETURN_IF(workspace(WS_BLAS3, (size_t)k * k * sizeof(double), &ws));
  double* o = static_cast<double*>(ws);
  const int64_t g = std::min<int64_t>((k * k + 255) / 256, 65536);
  if (sym)
    hipLaunchKernelGGL(expand_kernel<true>, dim3((unsigned)g), dim3(256), 0, stream(), k, a, rs,
                       cs, lower, unit, o);
  else
    hipLaunchKernelGGL(expand_kernel<false>, dim3((unsigned)g), dim3(256), 0, stream(), k, a, rs,
                       cs, lower, unit, o);
  DDLA_HIP_CHECK(hipGetLastError());
  *out = o;
  return 0;
}

// ------------------------------------------------------------------ device forms ---------
int gemmt_dev(int uploc, int transa, int transb, int64_t m, int64_t k, double alpha,
              const double* a, int64_t rsa, int64_t csa, const double* b, int64_t rsb, int64_t csb,
              double beta, double* c, int64_t rsc, int64_t csc) {
  if (m == 0) return 0;
  return gemm_strided(transa, transb, m, m, k, alpha, a, rsa, csa, b, rsb, csb, beta, c, rsc, csc,
                      tri_of(uploc));
}

// C = beta C + alpha op(A) op(B)^T + alpha op(B) op(A)^T (uplo triangle); op(B)^T is the view
// of B with its strides exchanged under the same trans flag
int syr2k_dev(int uploc, int transa, int transb, int64_t m, int64_t k, double alpha,
              const double* a, int64_t rsa, int64_t csa, const double* b, int64_t rsb,
              int64_t csb, double beta, double* c, int64_t rsc, int64_t csc) {
  if (m == 0) return 0;
  DDLA_RETURN_IF(gemm_strided(transa, transb, m, m, k, alpha, a, rsa, csa, b, csb, rsb, beta, c,
                              rsc, csc, tri_of(uploc)));
  return gemm_strided(transb, transa, m, m, k, alpha, b, rsb, csb, a, csa, rsa, 1.0, c, rsc, csc,
                      tri_of(uploc));
}

int symm_dev(int side, int uploa, int transb, int64_t m, int64_t n, double alpha, const double* a,
             int64_t rsa, int64_t csa, const double* b, int64_t rsb, int64_t csb, double beta,
             double* c, int64_t rsc, int64_t csc) {
  if (m == 0 || n == 0) return 0;
  const int64_t k = side == DDLA_LEFT ? m : n;
  double* S = nullptr;
  DDLA_RETURN_IF(expand(true, k, a, rsa, csa, uploa == DDLA_LOWER, false, &S));
  if (side == DDLA_LEFT)
    return gemm_strided(DDLA_NO_TRANSPOSE, transb, m, n, k, alpha, S, k, 1, b, rsb, csb, beta, c,
                        rsc, csc, 0);
  return gemm_strided(transb, DDLA_NO_TRANSPOSE, m, n, k, alpha, b, rsb, csb, S, k, 1, beta, c,
                      rsc, csc, 0);
}

int trmm3_dev(int side, int uploa, int transa, int diaga, int transb, int64_t m, int64_t n,
              double alpha, const double* a, int64_t rsa, int64_t csa, const double* b,
              int64_t rsb, int64_t csb, double beta, double* c, int64_t rsc, int64_t csc) {
  if (m == 0 || n == 0) return 0;
  const int64_t k = side == DDLA_LEFT ? m : n;
  double* T = nullptr;
  DDLA_RETURN_IF(expand(false, k, a, rsa, csa, uploa == DDLA_LOWER, diaga == DDLA_UNIT_DIAG, &T));
  if (side == DDLA_LEFT)
    return gemm_strided(transa, transb, m, n, k, alpha, T, k, 1, b, rsb, csb, beta, c, rsc, csc, 0);
  return gemm_strided(transb, transa, m, n, k, alpha, b, rsb, csb, T, k, 1, beta, c, rsc, csc, 0);
}

// B = alpha op(T) B (left) or alpha B op(T) (right): the product goes to a workspace, then
// over B
int trmm_dev(int side, int uploa, int transa, int diaga, int64_t m, int64_t n, double alpha,
             const double* a, int64_t rsa, int64_t csa, double* b, int64_t rsb, int64_t csb) {
  if (m == 0 || n == 0) return 0;
  void* ws = nullptr;
  DDLA_RETURN_IF(workspace(WS_BLAS3B, (size_t)m * n * sizeof(double), &ws));
  double* W = static_cast<double*>(ws);
  DDLA_RETURN_IF(trmm3_dev(side, uploa, transa, diaga, DDLA_NO_TRANSPOSE, m, n, alpha, a, rsa, csa,
                           b, rsb, csb, 0.0, W, n, 1));
  return launch_dscatter2d(m, n, W, n, b, rsb, csb, stream());
}

}  // namespace

#define API_INIT() DDLA_RETURN_IF(ensure_init())

extern "C" {

// ------------------------------------------------------------------------ GEMMT ---------
int ddla_dgemmt_dev(int uploc, int transa, int transb, ddla_dim_t m, ddla_dim_t k,
                    const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa,
                    const double* b, ddla_inc_t rsb, ddla_inc_t csb, const double* beta,
                    double* c, ddla_inc_t rsc, ddla_inc_t csc) {
  API_INIT();
  DDLA_ARG_CHECK(valid_uplo(uploc), "Matrix must be either upper or lower triangular.");
  DDLA_ARG_CHECK(valid_trans(transa) && valid_trans(transb), "bad trans flag");
  DDLA_ARG_CHECK(m >= 0 && k >= 0, "negative dimension");
  DDLA_ARG_CHECK(alpha && beta, "null alpha/beta");
  if (m == 0) return 0;
  DDLA_ARG_CHECK(c && (k == 0 || (a && b)), "null matrix pointer");
  return gemmt_dev(uploc, transa, transb, m, k, *alpha, a, rsa, csa, b, rsb, csb, *beta, c, rsc,
                   csc);
}

int ddla_dgemmt(int uploc, int transa, int transb, ddla_dim_t m, ddla_dim_t k, const double* alpha,
                const double* a, ddla_inc_t rsa, ddla_inc_t csa, const double* b, ddla_inc_t rsb,
                ddla_inc_t csb, const double* beta, double* c, ddla_inc_t rsc, ddla_inc_t csc) {
  API_INIT();
  DDLA_ARG_CHECK(valid_uplo(uploc), "Matrix must be either upper or lower triangular.");
  DDLA_ARG_CHECK(valid_trans(transa) && valid_trans(transb), "bad trans flag");
  DDLA_ARG_CHECK(m >= 0 && k >= 0, "negative dimension");
  DDLA_ARG_CHECK(alpha && beta, "null alpha/beta");
  if (m == 0) return 0;
  DDLA_ARG_CHECK(c && (k == 0 || (a && b)), "null matrix pointer");
  const bool ta = is_trans(transa), tb = is_trans(transb);
  DevBuf dA, dB, dC;
  DDLA_RETURN_IF(upload_view(ta ? k : m, ta ? m : k, a, rsa, csa, dA));
  DDLA_RETURN_IF(upload_view(tb ? m : k, tb ? k : m, b, rsb, csb, dB));
  DDLA_RETURN_IF(upload_view(m, m, c, rsc, csc, dC));  // the other triangle is written back as is
  DDLA_RETURN_IF(gemmt_dev(uploc, transa, transb, m, k, *alpha, dA.d(), ta ? m : k, 1, dB.d(),
                           tb ? k : m, 1, *beta, dC.d(), m, 1));
  return download_view(m, m, dC, c, rsc, csc);
}

// ------------------------------------------------------------------ SYRK / SYR2K ---------
int ddla_dsyrk(int uploc, int transa, ddla_dim_t m, ddla_dim_t k, const double* alpha,
               const double* a, ddla_inc_t rsa, ddla_inc_t csa, const double* beta, double* c,
               ddla_inc_t rsc, ddla_inc_t csc) {
  API_INIT();
  DDLA_ARG_CHECK(valid_uplo(uploc), "Matrix must be either upper or lower triangular.");
  DDLA_ARG_CHECK(valid_trans(transa), "bad trans flag");
  DDLA_ARG_CHECK(m >= 0 && k >= 0, "negative dimension");
  DDLA_ARG_CHECK(alpha && beta, "null alpha/beta");
  if (m == 0) return 0;
  DDLA_ARG_CHECK(c && (k == 0 || a), "null matrix pointer");
  const bool ta = is_trans(transa);
  const int64_t ar = ta ? k : m, ac = ta ? m : k;
  DevBuf dA, dC;
  DDLA_RETURN_IF(upload_view(ar, ac, a, rsa, csa, dA));
  DDLA_RETURN_IF(upload_view(m, m, c, rsc, csc, dC));
  // op(A) op(A)^T: the second operand is the transposed view of A under the same trans flag
  DDLA_RETURN_IF(gemmt_dev(uploc, transa, transa, m, k, *alpha, dA.d(), ac, 1, dA.d(), 1, ac,
                           *beta, dC.d(), m, 1));
  return download_view(m, m, dC, c, rsc, csc);
}

int ddla_dsyr2k_dev(int uploc, int transa, int transb, ddla_dim_t m, ddla_dim_t k,
                    const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa,
                    const double* b, ddla_inc_t rsb, ddla_inc_t csb, const double* beta,
                    double* c, ddla_inc_t rsc, ddla_inc_t csc) {
  API_INIT();
  DDLA_ARG_CHECK(valid_uplo(uploc), "Matrix must be either upper or lower triangular.");
  DDLA_ARG_CHECK(valid_trans(transa) && valid_trans(transb), "bad trans flag");
  DDLA_ARG_CHECK(m >= 0 && k >= 0, "negative dimension");
  DDLA_ARG_CHECK(alpha && beta, "null alpha/beta");
  if (m == 0) return 0;
  DDLA_ARG_CHECK(c && (k == 0 || (a && b)), "null matrix pointer");
  return syr2k_dev(uploc, transa, transb, m, k, *alpha, a, rsa, csa, b, rsb, csb, *beta, c, rsc,
                   csc);
}

int ddla_dsyr2k(int uploc, int transa, int transb, ddla_dim_t m, ddla_dim_t k, const double* alpha,
                const double* a, ddla_inc_t rsa, ddla_inc_t csa, const double* b, ddla_inc_t rsb,
                ddla_inc_t csb, const double* beta, double* c, ddla_inc_t rsc, ddla_inc_t csc) {
  API_INIT();
  DDLA_ARG_CHECK(valid_uplo(uploc), "Matrix must be either upper or lower triangular.");
  DDLA_ARG_CHECK(valid_trans(transa) && valid_trans(transb), "bad trans flag");
  DDLA_ARG_CHECK(m >= 0 && k >= 0, "negative dimension");
  DDLA_ARG_CHECK(alpha && beta, "null alpha/beta");
  if (m == 0) return 0;
  DDLA_ARG_CHECK(c && (k == 0 || (a && b)), "null matrix pointer");
  const bool ta = is_trans(transa), tb = is_trans(transb);
  DevBuf dA, dB, dC;
  DDLA_RETURN_IF(upload_view(ta ? k : m, ta ? m : k, a, rsa, csa, dA));
  DDLA_RETURN_IF(upload_view(tb ? k : m, tb ? m : k, b, rsb, csb, dB));
  DDLA_RETURN_IF(upload_view(m, m, c, rsc, csc, dC));
  DDLA_RETURN_IF(syr2k_dev(uploc, transa, transb, m, k, *alpha, dA.d(), ta ? m : k, 1, dB.d(),
                           tb ? m : k, 1, *beta, dC.d(), m, 1));
  return download_view(m, m, dC, c, rsc, csc);
}

// ------------------------------------------------------------------------- SYMM ---------
int ddla_dsymm_dev(int sidea, int uploa, int conja, int transb, ddla_dim_t m, ddla_dim_t n,
                   const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa,
                   const double* b, ddla_inc_t rsb, ddla_inc_t csb, const double* beta, double* c,
                   ddla_inc_t rsc, ddla_inc_t csc) {
  API_INIT();
  DDLA_ARG_CHECK(valid_side(sidea) && valid_uplo(uploa), "bad side / uplo flag");
  DDLA_ARG_CHECK(conja == 0 || conja == CONJ, "bad conj flag");
  DDLA_ARG_CHECK(valid_trans(transb), "bad trans flag");
  DDLA_ARG_CHECK(m >= 0 && n >= 0, "negative dimension");
  DDLA_ARG_CHECK(alpha && beta, "null alpha/beta");
  if (m == 0 || n == 0) return 0;
  DDLA_ARG_CHECK(a && b && c, "null matrix pointer");
  return symm_dev(sidea, uploa, transb, m, n, *alpha, a, rsa, csa, b, rsb, csb, *beta, c, rsc, csc);
}

int ddla_dsymm(int sidea, int uploa, int conja, int transb, ddla_dim_t m, ddla_dim_t n,
               const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa,
               const double* b, ddla_inc_t rsb, ddla_inc_t csb, const double* beta, double* c,
               ddla_inc_t rsc, ddla_inc_t csc) {
  API_INIT();
  DDLA_ARG_CHECK(valid_side(sidea) && valid_uplo(uploa), "bad side / uplo flag");
  DDLA_ARG_CHECK(conja == 0 || conja == CONJ, "bad conj flag");
  DDLA_ARG_CHECK(valid_trans(transb), "bad trans flag");
  DDLA_ARG_CHECK(m >= 0 && n >= 0, "negative dimension");
  DDLA_ARG_CHECK(alpha && beta, "null alpha/beta");
  if (m == 0 || n == 0) return 0;
  DDLA_ARG_CHECK(a && b && c, "null matrix pointer");
  const int64_t k = sidea == DDLA_LEFT ? m : n;
  const bool tb = is_trans(transb);
  DevBuf dA, dB, dC;
  DDLA_RETURN_IF(upload_view(k, k, a, rsa, csa, dA));
  DDLA_RETURN_IF(upload_view(tb ? n : m, tb ? m : n, b, rsb, csb, dB));
  DDLA_RETURN_IF(upload_view(m, n, c, rsc, csc, dC));
  DDLA_RETURN_IF(symm_dev(sidea, uploa, transb, m, n, *alpha, dA.d(), k, 1, dB.d(), tb ? m : n, 1,
                          *beta, dC.d(), n, 1));
  return download_view(m, n, dC, c, rsc, csc);
}

// ------------------------------------------------------------------ TRMM / TRMM3 ---------
int ddla_dtrmm_dev(int sidea, int uploa, int transa, int diaga, ddla_dim_t m, ddla_dim_t n,
                   const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa, double* b,
                   ddla_inc_t rsb, ddla_inc_t csb) {
  API_INIT();
  DDLA_ARG_CHECK(valid_side(sidea) && valid_uplo(uploa), "bad side / uplo flag");
  DDLA_ARG_CHECK(valid_trans(transa) && valid_diag(diaga), "bad trans / diag flag");
  DDLA_ARG_CHECK(m >= 0 && n >= 0, "negative dimension");
  DDLA_ARG_CHECK(alpha != nullptr, "null alpha");
  if (m == 0 || n == 0) return 0;
  DDLA_ARG_CHECK(a && b, "null matrix pointer");
  return trmm_dev(sidea, uploa, transa, diaga, m, n, *alpha, a, rsa, csa, b, rsb, csb);
}

int ddla_dtrmm(int sidea, int uploa, int transa, int diaga, ddla_dim_t m, ddla_dim_t n,
               const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa, double* b,
               ddla_inc_t rsb, ddla_inc_t csb) {
  API_INIT();
  DDLA_ARG_CHECK(valid_side(sidea) && valid_uplo(uploa), "bad side / uplo flag");
  DDLA_ARG_CHECK(valid_trans(transa) && valid_diag(diaga), "bad trans / diag flag");
  DDLA_ARG_CHECK(m >= 0 && n >= 0, "negative dimension");
  DDLA_ARG_CHECK(alpha != nullptr, "null alpha");
  if (m == 0 || n == 0) return 0;
  DDLA_ARG_CHECK(a && b, "null matrix pointer");
  const int64_t k = sidea == DDLA_LEFT ? m : n;
  DevBuf dA, dB;
  DDLA_RETURN_IF(upload_view(k, k, a, rsa, csa, dA));
  DDLA_RETURN_IF(upload_view(m, n, b, rsb, csb, dB));
  DDLA_RETURN_IF(trmm_dev(sidea, uploa, transa, diaga, m, n, *alpha, dA.d(), k, 1, dB.d(), n, 1));
  return download_view(m, n, dB, b, rsb, csb);
}

int ddla_dtrmm3_dev(int sidea, int uploa, int transa, int diaga, int transb, ddla_dim_t m,
                    ddla_dim_t n, const double* alpha, const double* a, ddla_inc_t rsa,
                    ddla_inc_t csa, const double* b, ddla_inc_t rsb, ddla_inc_t csb,
                    const double* beta, double* c, ddla_inc_t rsc, ddla_inc_t csc) {
  API_INIT();
  DDLA_ARG_CHECK(valid_side(sidea) && valid_uplo(uploa), "bad side / uplo flag");
  DDLA_ARG_CHECK(valid_trans(transa) && valid_trans(transb) && valid_diag(diaga),
                 "bad trans / diag flag");
  DDLA_ARG_CHECK(m >= 0 && n >= 0, "negative dimension");
  DDLA_ARG_CHECK(alpha && beta, "null alpha/beta");
  if (m == 0 || n == 0) return 0;
  DDLA_ARG_CHECK(a && b && c, "null matrix pointer");
  return trmm3_dev(sidea, uploa, transa, diaga, transb, m, n, *alpha, a, rsa, csa, b, rsb, csb,
                   *beta, c, rsc, csc);
}

int ddla_dtrmm3(int sidea, int uploa, int transa, int diaga, int transb, ddla_dim_t m,
                ddla_dim_t n, const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa,
                const double* b, ddla_inc_t rsb, ddla_inc_t csb, const double* beta, double* c,
                ddla_inc_t rsc, ddla_inc_t csc) {
  API_INIT();
  DDLA_ARG_CHECK(valid_side(sidea) && valid_uplo(uploa), "bad side / uplo flag");
  DDLA_ARG_CHECK(valid_trans(transa) && valid_trans(transb) && valid_diag(diaga),
                 "bad trans / diag flag");
  DDLA_ARG_CHECK(m >= 0 && n >= 0, "negative dimension");
  DDLA_ARG_CHECK(alpha && beta, "null alpha/beta");
  if (m == 0 || n == 0) return 0;
  DDLA_ARG_CHECK(a && b && c, "null matrix pointer");
  const int64_t k = sidea == DDLA_LEFT ? m : n;
  const bool tb = is_trans(transb);
  DevBuf dA, dB, dC;
  DDLA_RETURN_IF(upload_view(k, k, a, rsa, csa, dA));
  DDLA_RETURN_IF(upload_view(tb ? n : m, tb ? m : n, b, rsb, csb, dB));
  DDLA_RETURN_IF(upload_view(m, n, c, rsc, csc, dC));
  DDLA_RETURN_IF(trmm3_dev(sidea, uploa, transa, diaga, transb, m, n, *alpha, dA.d(), k, 1, dB.d(),
                           tb ? m : n, 1, *beta, dC.d(), n, 1));
  return download_view(m, n, dC, c, rsc, csc);
}

// ------------------------------------------------------------------------- TRSM ---------
int ddla_dtrsm(int sidea, int uploa, int transa, int diaga, ddla_dim_t m, ddla_dim_t n,
               const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa, double* b,
               ddla_inc_t rsb, ddla_inc_t csb) {
  API_INIT();
  DDLA_ARG_CHECK(valid_side(sidea) && valid_uplo(uploa), "bad side / uplo flag");
  DDLA_ARG_CHECK(valid_trans(transa) && valid_diag(diaga), "bad trans / diag flag");
  DDLA_ARG_CHECK(m >= 0 && n >= 0, "negative dimension");
  DDLA_ARG_CHECK(alpha != nullptr, "null alpha");
  if (m == 0 || n == 0) return 0;
  DDLA_ARG_CHECK(a && b, "null matrix pointer");
  const int64_t k = sidea == DDLA_LEFT ? m : n;
  DevBuf dA, dB;
  DDLA_RETURN_IF(upload_view(k, k, a, rsa, csa, dA));
  DDLA_RETURN_IF(upload_view(m, n, b, rsb, csb, dB));
  DDLA_RETURN_IF(trsm_rowmajor(sidea, uploa, transa, diaga, m, n, *alpha, dA.d(), k, dB.d(), n));
  return download_view(m, n, dB, b, rsb, csb);
}

}  // extern "C"
