// ddla_oracle.cpp -- CPU restatement of DDLA's managed FP64 dense path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP library and
// the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg may load it; the product library (ddla_amd/libddla_hip.so) never links,
// loads or calls it.
//
// What it restates (reference paths are relative to the DDLA repository root):
//   GEMM      DDLA/BLAS/Managed/3m.cs:13-49 (entry, beta scaling, early outs)
//             DDLA/BLAS/Managed/3m/GeMM.cs:14-73 (5-loop nest, MC=510 KC=144 NC=512,
//               row-major C handled by computing C^T = B^T A^T)
//             DDLA/BLAS/Managed/3m/Packing.cs:30-101 (MR/NR micro-panels, zero padded)
//             DDLA/BLAS/Managed/3m/MacroKernel.cs:24-97 (per-panel micro-kernel + add-back)
//             DDLA/Einsum/GEMMKernel.cs:36-240 (6x8 FMA tile, k ascending from 0.0)
//             DDLA/BLAS/Managed/1m.cs:86-155,251-307 (C += alpha*tile, C *= beta / C = 0)
//   AMAX      DDLA/BLAS/1v.cs:30-41 (first index of max |x|; NaN smallest)
//   LU        DDLA/Factorizations/LU.cs:133-465 (block 128, relative pivots)
//   Cholesky  DDLA/Factorizations/Cholesky.cs:52-236 (block 128, lower and upper)
//   TRSM      operation of DDLA/BLAS/3m.cs:287-314 (facade -> bli_dtrsm), blocked like
//             DDLA/BLAS/Managed/3m/TrSM.cs:62-195 (KC-row blocks, GEMM update + TRSV)
//   SYRK      operation of DDLA/BLAS/3m.cs:118-147 (facade -> bli_dsyrk)
//   TSLU      no reference counterpart: the FP32 tournament-pivoting panel of the mixed LU
//             (CALU, Grigori-Demmel-Xiang 2011), in the GPU kernel's operation order
//
// Bit-faithfulness: for alpha=1, beta=0 the GEMM result of every element is
//   c = (((0 + s_0) + s_1) + ... ) with s_p = fma-chain over k in KC-block p, from 0.0,
// exactly the managed GEMM's rounding.  The shipped LU/Cholesky call BLIS (closed) for
// TRSM/GEMM/SYRK/GER; their exact rounding is not reproducible here and the oracle uses
// the managed GEMM in their place ("parity unpinned" at the last-bit level for those
// calls; pinned by invariants and library cross-checks instead).
//
// Layout: every matrix argument is a strided view (pointer, row stride, column stride)
// exactly like DDLA's MatrixView (DDLA/Core/MatrixView.cs:22-31,483-485).

#include <immintrin.h>
#include <omp.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// BLIS-valued flags (DDLA/Misc/Flags/*.cs)
constexpr int kNoTrans = 0x00, kTrans = 0x08;
constexpr int kLower = 0xC0, kUpper = 0x60;
constexpr int kNonUnit = 0x000, kUnit = 0x100;
constexpr int kLeft = 0, kRight = 1;

int g_threads = 0;
int g_pivblock = 128;  // LU.PLUBlockSize / LUBlockSize (LU.cs:131,235)
int g_cholblock = 128; // Cholesky.BlockSize (Cholesky.cs:125)

int threads() {
  if (g_threads <= 0) {
    long np = sysconf(_SC_NPROCESSORS_ONLN);
    g_threads = std::max(1L, np / 2);  // ProcessorCount / 2 (MacroKernel.cs:72-73)
  }
  return g_threads;
}

struct MV {
  double* d;
  int64_t rows, cols, rs, cs;
  double& at(int64_t i, int64_t j) const { return d[i * rs + j * cs]; }
  MV T() const { return {d, cols, rows, cs, rs}; }
  MV sub(int64_t i, int64_t m, int64_t j, int64_t n) const {
    return {d + i * rs + j * cs, m, n, rs, cs};
  }
  bool empty() const { return rows == 0 || cols == 0; }
};

// ---------------------------------------------------------------------------------------
// Managed GEMM (GeMM.cs / Packing.cs / MacroKernel.cs / GEMMKernel.cs)
// ---------------------------------------------------------------------------------------
constexpr int64_t MC = 510, NC = 512, KC = 144, MR = 6, NR = 8;

inline int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// GEPack.Invoke (Packing.cs:60-85): panel i holds rows [i*blk, i*blk+blk) of src stored
// as [col][row-in-panel], rows past the end zero filled.
void pack_panels(const MV& src, double* buf, int64_t blk) {
  const int64_t rows = src.rows, cols = src.cols;
  const int64_t npan = align_up(rows, blk) / blk;
#pragma omp parallel for num_threads(threads()) schedule(static) if (npan >= 8)
  for (int64_t p = 0; p < npan; ++p) {
    double* dst = buf + p * cols * blk;
    const int64_t r0 = p * blk;
    const int64_t act = std::min(rows - r0, blk);
    for (int64_t c = 0; c < cols; ++c) {
      int64_t r = 0;
      for (; r < act; ++r) *dst++ = src.at(r0 + r, c);
      for (; r < blk; ++r) *dst++ = 0.0;
    }
  }
}

// GEMMKernel.MicroKernel (GEMMKernel.cs:36-240): 6x8 tile, twelve 4-wide accumulators
// starting at zero, acc = fma(a_r, b_c, acc) for k ascending; tile stored row-major.
inline void micro_kernel(int64_t k, const double* a, const double* b, double* c) {
  __m256d c00 = _mm256_setzero_pd(), c04 = _mm256_setzero_pd();
  __m256d c10 = _mm256_setzero_pd(), c14 = _mm256_setzero_pd();
  __m256d c20 = _mm256_setzero_pd(), c24 = _mm256_setzero_pd();
  __m256d c30 = _mm256_setzero_pd(), c34 = _mm256_setzero_pd();
  __m256d c40 = _mm256_setzero_pd(), c44 = _mm256_setzero_pd();
  __m256d c50 = _mm256_setzero_pd(), c54 = _mm256_setzero_pd();
  for (int64_t p = 0; p < k; ++p) {
    const __m256d b0 = _mm256_loadu_pd(b), b4 = _mm256_loadu_pd(b + 4);
    b += 8;
    __m256d av;
    av = _mm256_broadcast_sd(a + 0); c00 = _mm256_fmadd_pd(av, b0, c00); c04 = _mm256_fmadd_pd(av, b4, c04);
    av = _mm256_broadcast_sd(a + 1); c10 = _mm256_fmadd_pd(av, b0, c10); c14 = _mm256_fmadd_pd(av, b4, c14);
    av = _mm256_broadcast_sd(a + 2); c20 = _mm256_fmadd_pd(av, b0, c20); c24 = _mm256_fmadd_pd(av, b4, c24);
    av = _mm256_broadcast_sd(a + 3); c30 = _mm256_fmadd_pd(av, b0, c30); c34 = _mm256_fmadd_pd(av, b4, c34);
    av = _mm256_broadcast_sd(a + 4); c40 = _mm256_fmadd_pd(av, b0, c40); c44 = _mm256_fmadd_pd(av, b4, c44);
    av = _mm256_broadcast_sd(a + 5); c50 = _mm256_fmadd_pd(av, b0, c50); c54 = _mm256_fmadd_pd(av, b4, c54);
    a += 6;
  }
  _mm256_storeu_pd(c + 0, c00);  _mm256_storeu_pd(c + 4, c04);
  _mm256_storeu_pd(c + 8, c10);  _mm256_storeu_pd(c + 12, c14);
  _mm256_storeu_pd(c + 16, c20); _mm256_storeu_pd(c + 20, c24);
  _mm256_storeu_pd(c + 24, c30); _mm256_storeu_pd(c + 28, c34);
  _mm256_storeu_pd(c + 32, c40); _mm256_storeu_pd(c + 36, c44);
  _mm256_storeu_pd(c + 40, c50); _mm256_storeu_pd(c + 44, c54);
}

// Scal(beta, C) (Managed/1m.cs:251-307): beta==0 -> Set 0 (clears NaN), beta==1 -> no-op.
void scale_matrix(double beta, const MV& C) {
  if (beta == 1.0) return;
  for (int64_t i = 0; i < C.rows; ++i)
    for (int64_t j = 0; j < C.cols; ++j)
      C.at(i, j) = (beta == 0.0) ? 0.0 : C.at(i, j) * beta;
}

// GeMMInner (GeMM.cs:14-73) for C += alpha * A * B with A m x k, B k x n.
void gemm_inner(int64_t m, int64_t n, int64_t k, double alpha, MV A, MV B, MV C) {
  if (C.cs < C.rs) {  // preferCol == false: row-major C -> compute C^T = B^T A^T
    C = C.T();
    std::swap(m, n);
    MV a2 = B.T(), b2 = A.T();
    A = a2;
    B = b2;
  }
  const int64_t mce = std::min(MC, align_up(m, MR));
  const int64_t nce = std::min(NC, align_up(n, NR));
  std::vector<double> bufA(mce * KC), bufB(KC * nce);
  for (int64_t ic = 0; ic < m; ic += MC) {
    const int64_t mc = std::min(MC, m - ic);
    for (int64_t pc = 0; pc < k; pc += KC) {
      const int64_t kc = std::min(KC, k - pc);
      pack_panels(A.sub(ic, mc, pc, kc), bufA.data(), MR);
      for (int64_t jc = 0; jc < n; jc += NC) {
        const int64_t nc = std::min(NC, n - jc);
        pack_panels(B.sub(pc, kc, jc, nc).T(), bufB.data(), NR);
        const MV Cs = C.sub(ic, mc, jc, nc);
        const int64_t ipan = align_up(mc, MR) / MR;
        const int64_t ncal = align_up(nc, NR);
        // GEMMMacroKernel.Invoke (MacroKernel.cs:69-96), parallel over MR panels
#pragma omp parallel for num_threads(threads()) schedule(static) if (ipan >= 8)
        for (int64_t i = 0; i < ipan; ++i) {
          alignas(32) double tile[MR * NR];
          const double* pa = bufA.data() + i * kc * MR;
          const int64_t amr = std::min(mc - i * MR, MR);
          for (int64_t j = 0; j < ncal; j += NR) {
            const int64_t anr = std::min(nc - j, NR);
            micro_kernel(kc, pa, bufB.data() + j * kc, tile);
            // Add back: C += alpha * tile (Managed/1m.cs:86-155 -> fma(tile, alpha, C))
            for (int64_t r = 0; r < amr; ++r)
              for (int64_t c = 0; c < anr; ++c) {
                double& dst = Cs.at(i * MR + r, j + c);
                dst = std::fma(tile[r * NR + c], alpha, dst);
              }
          }
        }
      }
    }
  }
}

// Managed BlasProvider.GeMM (Managed/3m.cs:13-39) with op() via transposed views.
void gemm(int ta, int tb, double alpha, MV A, MV B, double beta, MV C) {
  if (ta == kTrans) A = A.T();
  if (tb == kTrans) B = B.T();
  const int64_t m = C.rows, n = C.cols, k = A.cols;
  if (m == 0 || n == 0) return;
  scale_matrix(beta, C);
  if (k == 0) return;
  gemm_inner(m, n, k, alpha, A, B, C);
}

// ---------------------------------------------------------------------------------------
// Level 1/2 pieces used by the factorizations
// ---------------------------------------------------------------------------------------

// AMax (DDLA/BLAS/1v.cs:30-41): index of the first maximum of |x|; NaN counts as
// smaller than every number; empty -> 0.
int64_t amax(int64_t n, const double* x, int64_t inc) {
  int64_t best = 0;
  double bv = -1.0;  // below every |x|; NaN never compares greater
  for (int64_t i = 0; i < n; ++i) {
    const double v = std::fabs(x[i * inc]);
    if (v > bv) { bv = v; best = i; }
  }
  return best;
}

void swap_rows(const MV& X, int64_t i, int64_t j) {
  for (int64_t c = 0; c < X.cols; ++c) std::swap(X.at(i, c), X.at(j, c));
}

// ApplyPivInternel (LU.cs:376-429): sequential interchanges, relative offsets, p>0 only.
void apply_piv(int side, bool forward, const MV& X, const int32_t* p, int64_t np) {
  if (X.rows == 0 || X.cols == 0) return;
  const MV Y = (side == kLeft) ? X : X.T();
  if (forward) {
    for (int64_t i = 0; i < np; ++i)
      if (p[i] > 0) swap_rows(Y, i, i + p[i]);
  } else {
    for (int64_t i = np - 1; i >= 0; --i)
      if (p[i] > 0) swap_rows(Y, i, i + p[i]);
  }
}

// ---------------------------------------------------------------------------------------
// TRSM: op(A) X = alpha B (Left) or X op(A) = alpha B (Right); B overwritten with X.
// Reduced to a left, lower-or-upper, no-trans solve on a transposed view, then blocked
// by KC rows like TrSMLeftLower/TrSMLeftUpper (Managed/3m/TrSM.cs:62-121): GEMM update of
// the block from the solved part, then a column-wise substitution inside the block.
// ---------------------------------------------------------------------------------------
void trsm(int side, int uplo, int trans, int diag, double alpha, MV A, MV B) {
  if (B.rows == 0 || B.cols == 0) return;
  scale_matrix(alpha, B);
  bool lower = (uplo == kLower);
  if (trans == kTrans) { A = A.T(); lower = !lower; }
  if (side == kRight) {  // X A' = B  <=>  A'^T X^T = B^T
    A = A.T();
    lower = !lower;
    B = B.T();
  }
  const bool unit = (diag == kUnit);
  const int64_t m = B.rows, nrhs = B.cols;
  const int64_t blk = KC;
  auto solve_block = [&](int64_t i0, int64_t bs) {
    const MV Akk = A.sub(i0, bs, i0, bs);
#pragma omp parallel for num_threads(threads()) schedule(static) if (nrhs >= 64)
    for (int64_t c = 0; c < nrhs; ++c) {
      if (lower) {
        for (int64_t i = 0; i < bs; ++i) {
          double s = B.at(i0 + i, c);
          for (int64_t j = 0; j < i; ++j) s -= Akk.at(i, j) * B.at(i0 + j, c);
          B.at(i0 + i, c) = unit ? s : s / Akk.at(i, i);
        }
      } else {
        for (int64_t i = bs - 1; i >= 0; --i) {
          double s = B.at(i0 + i, c);
          for (int64_t j = i + 1; j < bs; ++j) s -= Akk.at(i, j) * B.at(i0 + j, c);
          B.at(i0 + i, c) = unit ? s : s / Akk.at(i, i);
        }
      }
    }
  };
  if (lower) {
    for (int64_t i0 = 0; i0 < m; i0 += blk) {
      const int64_t bs = std::min(blk, m - i0);
      if (i0 > 0)
        gemm(kNoTrans, kNoTrans, -1.0, A.sub(i0, bs, 0, i0), B.sub(0, i0, 0, nrhs), 1.0,
             B.sub(i0, bs, 0, nrhs));
      solve_block(i0, bs);
    }
  } else {
    int64_t end = m;
    while (end > 0) {
      const int64_t bs = std::min(blk, end);
      const int64_t i0 = end - bs;
      if (end < m)
        gemm(kNoTrans, kNoTrans, -1.0, A.sub(i0, bs, end, m - end), B.sub(end, m - end, 0, nrhs),
             1.0, B.sub(i0, bs, 0, nrhs));
      solve_block(i0, bs);
      end = i0;
    }
  }
}

// SYRK on one triangle: C = beta*C + alpha*op(A)*op(A)^T, only the uplo triangle written.
void syrk(int uplo, int trans, double alpha, MV A, double beta, MV C) {
  if (trans == kTrans) A = A.T();
  const int64_t n = C.rows, k = A.cols;
  if (n == 0) return;
  const bool lower = (uplo == kLower);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < n; ++j)
      if (lower ? (j <= i) : (j >= i))
        C.at(i, j) = (beta == 0.0) ? 0.0 : (beta == 1.0 ? C.at(i, j) : C.at(i, j) * beta);
  if (k == 0) return;
  const int64_t blk = 128;
  std::vector<double> tmp(blk * blk);
  for (int64_t i0 = 0; i0 < n; i0 += blk) {
    const int64_t bs = std::min(blk, n - i0);
    // off-diagonal rectangle
    if (lower && i0 > 0)
      gemm(kNoTrans, kTrans, alpha, A.sub(i0, bs, 0, k), A.sub(0, i0, 0, k), 1.0,
           C.sub(i0, bs, 0, i0));
    if (!lower && i0 + bs < n)
      gemm(kNoTrans, kTrans, alpha, A.sub(i0, bs, 0, k), A.sub(i0 + bs, n - i0 - bs, 0, k),
           1.0, C.sub(i0, bs, i0 + bs, n - i0 - bs));
    // diagonal block through a scratch tile, triangle added back
    MV T{tmp.data(), bs, bs, bs, 1};
    gemm(kNoTrans, kTrans, alpha, A.sub(i0, bs, 0, k), A.sub(i0, bs, 0, k), 0.0, T);
    for (int64_t i = 0; i < bs; ++i)
      for (int64_t j = 0; j < bs; ++j)
        if (lower ? (j <= i) : (j >= i)) C.at(i0 + i, i0 + j) += T.at(i, j);
  }
}

// ---------------------------------------------------------------------------------------
// LU with partial pivoting (LU.cs:251-465)
// ---------------------------------------------------------------------------------------

// PLUDecUnblock + GetAndApplyPivUnblock (LU.cs:344-374). Returns 0 or (i+1) at the first
// zero pivot (the reference throws ArgumentOutOfRangeException from InvScal, 1v.cs:99).
int64_t plu_unblock(MV A, int32_t* piv) {
  const int64_t len = std::min(A.rows, A.cols);
  MV full = A;
  for (int64_t i = 0; i < len; ++i) {
    const int64_t ip = amax(A.rows, &A.at(0, 0), A.rs);
    if (ip != 0) swap_rows(full, 0, ip);
    piv[i] = static_cast<int32_t>(ip);
    const double a11 = A.at(0, 0);
    if (a11 == 0.0) return i + 1;
    const double r = 1.0 / a11;  // InvScal: Scal by 1/alpha (1v.cs:102)
    for (int64_t rr = 1; rr < A.rows; ++rr) A.at(rr, 0) *= r;
    // GeR(-1, a21, a12t, A22): A22 -= a21 a12t
    const int64_t mr = A.rows - 1, nc = A.cols - 1;
    if (mr > 0 && nc > 0) {
#pragma omp parallel for num_threads(threads()) schedule(static) if (mr * nc >= 65536)
      for (int64_t rr = 1; rr <= mr; ++rr) {
        const double l = A.at(rr, 0);
        for (int64_t c = 1; c <= nc; ++c) A.at(rr, c) = std::fma(-l, A.at(0, c), A.at(rr, c));
      }
    }
    A = A.sub(1, A.rows - 1, 1, A.cols - 1);
    full = full.sub(1, full.rows - 1, 0, full.cols);
  }
  return 0;
}

// PLUDecBlock (LU.cs:289-342)
int64_t plu_block(MV A, int32_t* p) {
  const int64_t n = A.rows;
  for (int64_t j = 0; j < n; j += g_pivblock) {
    const int64_t b = std::min<int64_t>(g_pivblock, n - j);
    int32_t* pc = p + j;
    const int64_t info = plu_unblock(A.sub(j, n - j, j, b), pc);
    if (info) return j + info;
    apply_piv(kLeft, true, A.sub(j, n - j, 0, j), pc, b);
    apply_piv(kLeft, true, A.sub(j, n - j, j + b, n - j - b), pc, b);
    const MV a11 = A.sub(j, b, j, b), a12 = A.sub(j, b, j + b, n - j - b);
    const MV a21 = A.sub(j + b, n - j - b, j, b), A22 = A.sub(j + b, n - j - b, j + b, n - j - b);
    trsm(kLeft, kLower, kNoTrans, kUnit, 1.0, a11, a12);
    gemm(kNoTrans, kNoTrans, -1.0, a21, a12, 1.0, A22);
  }
  return 0;
}

// LU without pivoting (LU.cs:169-203). Returns 0 or (i+1) at the first zero pivot.
int64_t lu_unblock(MV A) {
  const int64_t len = std::min(A.rows, A.cols);
  for (int64_t i = 1; i < len; ++i) {
    const double a11 = A.at(0, 0);
    if (a11 == 0.0) return i;
    const double r = 1.0 / a11;
    for (int64_t rr = 1; rr < A.rows; ++rr) A.at(rr, 0) *= r;
    for (int64_t rr = 1; rr < A.rows; ++rr) {
      const double l = A.at(rr, 0);
      for (int64_t c = 1; c < A.cols; ++c) A.at(rr, c) = std::fma(-l, A.at(0, c), A.at(rr, c));
    }
    A = A.sub(1, A.rows - 1, 1, A.cols - 1);
  }
  return 0;
}

int64_t lu_block(MV A) {
  const int64_t n = A.rows;
  for (int64_t j = 0; j < n; j += g_pivblock) {
    const int64_t b = std::min<int64_t>(g_pivblock, n - j);
    const MV a11 = A.sub(j, b, j, b), a12 = A.sub(j, b, j + b, n - j - b);
    const MV a21 = A.sub(j + b, n - j - b, j, b), A22 = A.sub(j + b, n - j - b, j + b, n - j - b);
    const int64_t info = lu_unblock(a11);
    if (info) return j + info;
    trsm(kLeft, kLower, kNoTrans, kUnit, 1.0, a11, a12);
    trsm(kRight, kUpper, kNoTrans, kNonUnit, 1.0, a11, a21);
    gemm(kNoTrans, kNoTrans, -1.0, a21, a12, 1.0, A22);
  }
  return 0;
}

// ---------------------------------------------------------------------------------------
// Cholesky (Cholesky.cs:52-236)
// ---------------------------------------------------------------------------------------

// CholeskyLowerUnblock (Cholesky.cs:127-140) with Sqrt (:228-236): a pivot that is not
// > 0 is reported as info (finite <= 0 throws in the reference; NaN and -inf make the
// following InvScal throw there, Managed/1v.cs:283-292).
int64_t chol_lower_unblock(MV A) {
  const int64_t n = A.rows;
  for (int64_t i = 0; i < n; ++i) {
    double a = A.at(i, i);
    if (!(a > 0.0)) return i + 1;
    a = std::sqrt(a);
    A.at(i, i) = a;
    if (i + 1 < n) {
      if (a != 1.0) {
        const double r = 1.0 / a;
        for (int64_t rr = i + 1; rr < n; ++rr) A.at(rr, i) *= r;
      }
      // Rank1(Lower, -1, a21): A22[r][c] += a21[c] * (-a21[r]) for c <= r
      for (int64_t rr = i + 1; rr < n; ++rr) {
        const double xr = -A.at(rr, i);
        for (int64_t c = i + 1; c <= rr; ++c) A.at(rr, c) = std::fma(A.at(c, i), xr, A.at(rr, c));
      }
    }
  }
  return 0;
}

// CholeskyUpperUnblock (Cholesky.cs:162-186): row-oriented, left-looking.
int64_t chol_upper_unblock(MV A) {
  const int64_t n = A.rows;
  for (int64_t i = 0; i < n; ++i) {
    double a = A.at(i, i);
    for (int64_t r = 0; r < i; ++r) a -= A.at(r, i) * A.at(r, i);
    if (!(a > 0.0)) return i + 1;
    a = std::sqrt(a);
    A.at(i, i) = a;
    for (int64_t c = i + 1; c < n; ++c) {
      double s = A.at(i, c);
      for (int64_t r = 0; r < i; ++r) s -= A.at(r, i) * A.at(r, c);
      A.at(i, c) = s;
    }
    if (i + 1 < n && a != 1.0) {
      const double r = 1.0 / a;
      for (int64_t c = i + 1; c < n; ++c) A.at(i, c) *= r;
    }
  }
  return 0;
}

int64_t chol_lower_block(MV A) {
  const int64_t n = A.rows;
  for (int64_t j = 0; j < n; j += g_cholblock) {
    const int64_t b = std::min<int64_t>(g_cholblock, n - j);
    const MV a11 = A.sub(j, b, j, b), a21 = A.sub(j + b, n - j - b, j, b);
    const MV A22 = A.sub(j + b, n - j - b, j + b, n - j - b);
    const int64_t info = chol_lower_unblock(a11);
    if (info) return j + info;
    trsm(kRight, kLower, kTrans, kNonUnit, 1.0, a11, a21);
    syrk(kLower, kNoTrans, -1.0, a21, 1.0, A22);
  }
  return 0;
}

int64_t chol_upper_block(MV A) {
  const int64_t n = A.rows;
  for (int64_t j = 0; j < n; j += g_cholblock) {
    const int64_t b = std::min<int64_t>(g_cholblock, n - j);
    const MV a11 = A.sub(j, b, j, b), a12 = A.sub(j, b, j + b, n - j - b);
    const MV A22 = A.sub(j + b, n - j - b, j + b, n - j - b);
    const int64_t info = chol_upper_unblock(a11);
    if (info) return j + info;
    trsm(kLeft, kUpper, kTrans, kNonUnit, 1.0, a11, a12);
    syrk(kUpper, kNoTrans, -1.0, a12.T(), 1.0, A22);
  }
  return 0;
}

uint64_t splitmix(uint64_t x) {
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}

// ------------------------------------------------------------ tournament pivoting -------
// No counterpart in the reference (its LU pivots column by column, LU.cs:344-374): the FP32
// tournament-pivoting panel of the mixed-precision LU, i.e. CALU's tournament (Grigori,
// Demmel, Xiang, SIAM J. Matrix Anal. Appl. 32, 2011) over stacks of 256 rows, restated
// serially.  The elimination of a stack follows the GPU kernel's operation order (columns in
// blocks of 16; each block's U12 by forward substitution over the block's pivot rows, then
// the rank-16 update of the remaining rows; every update one fmaf, index ascending; the
// multiplier is the value times the pivot's reciprocal), so selection and factored rows can
// be compared bit for bit.
namespace tourn {

constexpr int ROWS = 256, NB = 128, CB = 16;

// GEPP on the cnt x NB row-major stack S in place (rows not moved).  order: stack rows in pivot
// order (min(cnt, NB) of them).  Returns the first zero-pivot column + 1, or 0.
int stack_gepp(std::vector<float>& S, int cnt, std::vector<int>& order) {
  order.clear();
  if (cnt <= 0) return 0;
  const int npiv = std::min(cnt, NB);
  order.assign(npiv, 0);
  std::vector<char> active(cnt, 1);
  int zero = 0;
  auto at = [&](int r, int c) -> float& { return S[(size_t)r * NB + c]; };
  for (int j0 = 0; j0 < npiv; j0 += CB) {
    for (int i = 0; i < CB && j0 + i < npiv; ++i) {
      const int j = j0 + i;
      float best = -2.f;
      int p = -1;
      for (int t = 0; t < cnt; ++t) {  // largest key, first row among equal keys; NaN -> -1
        const float key = active[t] ? std::fmax(std::fabs(at(t, j)), -1.f) : -2.f;
        if (p < 0 || key > best) {
          best = key;
          p = t;
        }
      }
      order[j] = p;
      active[p] = 0;
      if (!(best > 0.f) && zero == 0) zero = j + 1;
      const float piv = at(p, j);
      const float rinv = piv != 0.f ? 1.f / piv : 0.f;
      for (int t = 0; t < cnt; ++t) {
        if (!active[t]) continue;
        const float l = at(t, j) * rinv;
        at(t, j) = l;
        for (int c = j + 1; c < j0 + CB; ++c) at(t, c) = std::fmaf(-l, at(p, c), at(t, c));
      }
    }
    const int c1 = j0 + CB;
    if (c1 >= npiv) break;
    std::vector<float> U12((size_t)(NB - c1) * CB);
    for (int c = c1; c < NB; ++c) {
      float u[CB];
      for (int i = 0; i < CB; ++i) {
        const int pr = order[j0 + i];
        float x = at(pr, c);
        for (int k = 0; k < i; ++k) x = std::fmaf(-at(pr, j0 + k), u[k], x);
        u[i] = x;
        at(pr, c) = x;
        U12[(size_t)(c - c1) * CB + i] = x;
      }
    }
    for (int t = 0; t < cnt; ++t) {
      if (!active[t]) continue;
      for (int c = c1; c < NB; ++c) {
        float x = at(t, c);
        for (int i = 0; i < CB; ++i) x = std::fmaf(-at(t, j0 + i), U12[(size_t)(c - c1) * CB + i], x);
        at(t, c) = x;
      }
    }
  }
  return zero;
}

// Stack of the panel rows `rows` (original values), eliminated; winners (panel rows, pivot
// order) returned, factored stack left in S.
std::vector<int> reduce(const float* P, int64_t ld, const std::vector<int>& rows,
                        std::vector<float>& S, std::vector<int>& order, int* zero) {
  const int cnt = (int)rows.size();
  S.assign((size_t)cnt * NB, 0.f);
  for (int r = 0; r < cnt; ++r)
    for (int c = 0; c < NB; ++c) S[(size_t)r * NB + c] = P[(int64_t)rows[r] * ld + c];
  *zero = stack_gepp(S, cnt, order);
  std::vector<int> win(order.size());
  for (size_t i = 0; i < order.size(); ++i) win[i] = rows[order[i]];
  return win;
}

}  // namespace tourn

}  // namespace

extern "C" {

int orc_set_threads(int t) { g_threads = t; return threads(); }
int orc_get_threads(void) { return threads(); }
void orc_set_block(int pivblock, int cholblock) {
  if (pivblock > 0) g_pivblock = pivblock;
  if (cholblock > 0) g_cholblock = cholblock;
}

// x[i] = 2*u_i - 1 with u_i the top 53 bits of splitmix64(seed + (i+1)*golden) / 2^53.
void orc_fill_uniform(double* x, int64_t n, uint64_t seed) {
#pragma omp parallel for num_threads(threads()) schedule(static) if (n >= (1 << 16))
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t z = splitmix(seed + static_cast<uint64_t>(i + 1) * 0x9E3779B97F4A7C15ull);
    x[i] = static_cast<double>(z >> 11) * 0x1.0p-52 - 1.0;
  }
}

int orc_dgemm(int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
              int64_t rsa, int64_t csa, const double* b, int64_t rsb, int64_t csb, double beta,
              double* c, int64_t rsc, int64_t csc) {
  MV A{const_cast<double*>(a), ta == kTrans ? k : m, ta == kTrans ? m : k, rsa, csa};
  MV B{const_cast<double*>(b), tb == kTrans ? n : k, tb == kTrans ? k : n, rsb, csb};
  MV C{c, m, n, rsc, csc};
  gemm(ta, tb, alpha, A, B, beta, C);
  return 0;
}

int64_t orc_damaxv(int64_t n, const double* x, int64_t inc) { return amax(n, x, inc); }

int orc_dtrsm(int side, int uplo, int trans, int diag, int64_t m, int64_t n, double alpha,
              const double* a, int64_t rsa, int64_t csa, double* b, int64_t rsb, int64_t csb) {
  const int64_t ka = (side == kLeft) ? m : n;
  MV A{const_cast<double*>(a), ka, ka, rsa, csa};
  MV B{b, m, n, rsb, csb};
  trsm(side, uplo, trans, diag, alpha, A, B);
  return 0;
}

int orc_dsyrk(int uplo, int trans, int64_t m, int64_t k, double alpha, const double* a,
              int64_t rsa, int64_t csa, double beta, double* c, int64_t rsc, int64_t csc) {
  MV A{const_cast<double*>(a), trans == kTrans ? k : m, trans == kTrans ? m : k, rsa, csa};
  MV C{c, m, m, rsc, csc};
  syrk(uplo, trans, alpha, A, beta, C);
  return 0;
}

// PLUDecompose (LU.cs:251-264): blocked iff n >= block.
int64_t orc_dgetrf(int64_t n, double* a, int64_t rs, int64_t cs, int32_t* piv) {
  MV A{a, n, n, rs, cs};
  if (n == 0) return 0;
  return (n >= g_pivblock) ? plu_block(A, piv) : plu_unblock(A, piv);
}

// PLU panel only (PLUDecUnblock on an m x b panel); used to pin the panel kernel.
int64_t orc_dgetf2(int64_t m, int64_t nb, double* a, int64_t rs, int64_t cs, int32_t* piv) {
  MV A{a, m, nb, rs, cs};
  return plu_unblock(A, piv);
}

// PLUDecUnblock (LU.cs:344-374) in FP32 on an m x nb panel (element (i, j) at a[i*rs + j*cs]):
// the same steps as plu_unblock in single precision -- AMax's first index of the largest |x|
// (NaN below every number), a physical row swap, the column scaled by the reciprocal of the
// pivot (InvScal, 1v.cs:102), then GeR as one fused multiply-add per element.  The FP32 panel
// of the mixed-precision LU has no counterpart in the reference; this pins its arithmetic.
int64_t orc_sgetf2(int64_t m, int64_t nb, float* a, int64_t rs, int64_t cs, int32_t* piv) {
  auto at = [&](int64_t i, int64_t j) -> float& { return a[i * rs + j * cs]; };
  const int64_t len = std::min(m, nb);
  for (int64_t i = 0; i < len; ++i) {
    int64_t ip = i;
    float bv = -1.0f;
    for (int64_t r = i; r < m; ++r) {
      const float v = std::fabs(at(r, i));
      if (v > bv) { bv = v; ip = r; }
    }
    if (ip != i)
      for (int64_t c = 0; c < nb; ++c) std::swap(at(i, c), at(ip, c));
    piv[i] = static_cast<int32_t>(ip - i);
    const float a11 = at(i, i);
    if (a11 == 0.0f) return i + 1;
    const float r = 1.0f / a11;
    for (int64_t rr = i + 1; rr < m; ++rr) at(rr, i) *= r;
    for (int64_t rr = i + 1; rr < m; ++rr) {
      const float l = at(rr, i);
      for (int64_t c = i + 1; c < nb; ++c) at(rr, c) = std::fmaf(-l, at(i, c), at(rr, c));
    }
  }
  return 0;
}

// LUDecompose without pivoting (LU.cs:133-142): blocked iff n > block.
int64_t orc_dgetrf_nopiv(int64_t n, double* a, int64_t rs, int64_t cs) {
  MV A{a, n, n, rs, cs};
  if (n == 0) return 0;
  return (n <= g_pivblock) ? lu_unblock(A) : lu_block(A);
}

// ApplyPiv (LU.cs:237-249)
int orc_dlaswp(int side, int forward, int64_t m, int64_t n, double* x, int64_t rs, int64_t cs,
               const int32_t* piv, int64_t np) {
  MV X{x, m, n, rs, cs};
  apply_piv(side, forward != 0, X, piv, np);
  return 0;
}

// PLUSolve (LU.cs:266-287, 431-465)
int orc_dgetrs(int trans, int64_t n, int64_t nrhs, const double* a, int64_t rsa, int64_t csa,
               const int32_t* piv, double* b, int64_t rsb, int64_t csb) {
  MV A{const_cast<double*>(a), n, n, rsa, csa};
  MV B{b, n, nrhs, rsb, csb};
  if (trans == kTrans) {
    trsm(kLeft, kUpper, kTrans, kNonUnit, 1.0, A, B);
    trsm(kLeft, kLower, kTrans, kUnit, 1.0, A, B);
    apply_piv(kLeft, false, B, piv, n);
  } else {
    apply_piv(kLeft, true, B, piv, n);
    trsm(kLeft, kLower, kNoTrans, kUnit, 1.0, A, B);
    trsm(kLeft, kUpper, kNoTrans, kNonUnit, 1.0, A, B);
  }
  return 0;
}

// CholDecompose (Cholesky.cs:52-69): blocked iff n > block.
int64_t orc_dpotrf(int uplo, int64_t n, double* a, int64_t rs, int64_t cs) {
  MV A{a, n, n, rs, cs};
  if (n == 0) return 0;
  if (uplo == kLower) return (n > g_cholblock) ? chol_lower_block(A) : chol_lower_unblock(A);
  return (n > g_cholblock) ? chol_upper_block(A) : chol_upper_unblock(A);
}

// CholeskySolveUncheck (Cholesky.cs:210-226)
int orc_dpotrs(int uplo, int64_t n, int64_t nrhs, const double* a, int64_t rsa, int64_t csa,
               double* b, int64_t rsb, int64_t csb) {
  MV A{const_cast<double*>(a), n, n, rsa, csa};
  MV B{b, n, nrhs, rsb, csb};
  if (uplo == kLower) {
    trsm(kLeft, kLower, kNoTrans, kNonUnit, 1.0, A, B);
    trsm(kLeft, kUpper, kNoTrans, kNonUnit, 1.0, A.T(), B);
  } else {
    trsm(kLeft, kLower, kNoTrans, kNonUnit, 1.0, A.T(), B);
    trsm(kLeft, kUpper, kNoTrans, kNonUnit, 1.0, A, B);
  }
  return 0;
}

// Tournament-pivoting panel (FP32, m >= 128 rows of 128 columns, row-major, ld): leaves of
// 256 consecutive rows, a binary tree over them (left set stacked above right set; the last
// node of an odd level moves up unchanged).  Outputs: sel[128] = panel rows in pivot order,
// top[128 x 128] = their factored rows (L11 \ U11), the (dst, src) row list of the panel's
// interchanges (selected rows to the top; displaced top rows, ascending, into the vacated rows
// in pivot order) and its length.  Returns the root's first zero-pivot column + 1, or 0.
int orc_stslu(int64_t m, const float* P, int64_t ld, int32_t* sel, float* top, int32_t* list,
              int32_t* count) {
  using namespace tourn;
  if (m < NB) return -1;
  const int leaves = (int)((m + ROWS - 1) / ROWS);
  std::vector<std::vector<int>> sets(leaves);
  std::vector<float> S;
  std::vector<int> order;
  int zero = 0;
  for (int l = 0; l < leaves; ++l) {
    std::vector<int> rows;
    for (int64_t r = (int64_t)l * ROWS; r < std::min<int64_t>(m, (int64_t)(l + 1) * ROWS); ++r)
      rows.push_back((int)r);
    sets[l] = reduce(P, ld, rows, S, order, &zero);
  }
  while (sets.size() > 1) {
    std::vector<std::vector<int>> up;
    for (size_t i = 0; i < sets.size(); i += 2) {
      if (i + 1 == sets.size()) {
        up.push_back(sets[i]);
        continue;
      }
      std::vector<int> rows = sets[i];
      rows.insert(rows.end(), sets[i + 1].begin(), sets[i + 1].end());
      up.push_back(reduce(P, ld, rows, S, order, &zero));
    }
    sets.swap(up);
  }
  // S / order / zero are the root's (the last reduction performed)
  const std::vector<int>& win = sets[0];
  for (int i = 0; i < NB; ++i) {
    sel[i] = win[i];
    for (int c = 0; c < NB; ++c) top[(size_t)i * NB + c] = S[(size_t)order[i] * NB + c];
  }
  std::vector<char> is_top_sel(NB, 0);
  for (int i = 0; i < NB; ++i)
    if (win[i] < NB) is_top_sel[win[i]] = 1;
  std::vector<int> disp, vac;
  for (int p = 0; p < NB; ++p)
    if (!is_top_sel[p]) disp.push_back(p);
  for (int i = 0; i < NB; ++i)
    if (win[i] >= NB) vac.push_back(win[i]);
  int cnt = 0;
  for (int i = 0; i < NB; ++i)
    if (win[i] != i) {
      list[2 * cnt] = i;
      list[2 * cnt + 1] = win[i];
      ++cnt;
    }
  for (size_t k = 0; k < disp.size(); ++k) {
    list[2 * cnt] = vac[k];
    list[2 * cnt + 1] = disp[k];
    ++cnt;
  }
  *count = cnt;
  return zero;
}

}  // extern "C"
