// FP64 factorization drivers (host side): blocked PLU, solves, blocked Cholesky.
//
//   dgetrf_internal  PLUDecBlock (DDLA/Factorizations/LU.cs:289-342): per 128-column panel
//                    panel factorization with pivoting, interchanges on the columns left
//                    and right of the panel, A12 <- L11^-1 A12, A22 <- A22 - A21 A12.
//   dgetrs_internal  PLUSolveInternel (LU.cs:431-465).
//   dpotrf_internal  CholeskyLowerBlock (DDLA/Factorizations/Cholesky.cs:142-160); the upper
//                    form factors the transposed storage with the same kernels.
//   dpotrs_internal  CholeskySolveUncheck (Cholesky.cs:210-226).
//   dldlt_internal   LDLTLowerBlock (DDLA/Factorizations/LDLT.cs:160-202); the upper form
//                    (LDLTUpperBlock, :223-265) factors the transposed storage.
//   dldlt_solve_internal  LDLTSolveUncheck (LDLT.cs:267-295).
#include <algorithm>

#include "ddla_internal.hpp"

namespace ddla {

namespace {
constexpr int NBLK = 128;  // LU.PLUBlockSize (LU.cs:235), Cholesky.BlockSize (Cholesky.cs:125)

// op(A) X = B for the stored triangle `lower` of A: the persistent solver (trsv.hip) for up
// to 8 right-hand sides, the blocked TRSM otherwise
int tri_solve(bool lower, bool unit, bool trans, int64_t n, int64_t nrhs, const double* a,
              int64_t lda, double* b, int64_t ldb) {
  hipStream_t s = stream();
  if (nrhs <= 8) {
    void* ws = nullptr;
    DDLA_RETURN_IF(workspace(WS_DINV, trsv_dinv_elems(n) * sizeof(double), &ws));
    double* dinv = static_cast<double*>(ws);
    const bool op_lower = trans ? !lower : lower;
    DDLA_RETURN_IF(trsv_prepare<double>(op_lower, unit, trans, n, a, lda, dinv, s));
    return trsv_solve<double>(op_lower, trans, n, (int)nrhs, a, lda, dinv, b, ldb, s);
  }
  return trsm_left<double>(lower, unit, trans, n, nrhs, a, lda, b, ldb, s);
}

// after the persistent triangular solves (<= 8 right-hand sides): did a wait give up?
int solves_done(int64_t nrhs, hipStream_t s) {
  if (nrhs > 8) return 0;
  return check_abort(trsv_abort_word(), s, "triangular solve");
}

int read_info(int32_t* dinfo, int32_t* info_host) {
  DDLA_HIP_CHECK(hipMemcpyAsync(info_host, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, stream()));
  DDLA_HIP_CHECK(hipStreamSynchronize(stream()));
  return 0;
}
}  // namespace

// Workspace layout of the pivot slot: [info word][counts: npanels][lists: npanels*4*NBLK]
int lu_bookkeeping(int64_t n, int32_t** dinfo, int** counts, int** lists) {
  const int64_t np = (n + NBLK - 1) / NBLK;
  void* ws = nullptr;
  const size_t bytes = 64 + np * sizeof(int) + np * 4 * NBLK * sizeof(int) + (size_t)n * sizeof(int);
  DDLA_RETURN_IF(workspace(WS_PIVOT, bytes, &ws));
  char* p = static_cast<char*>(ws);
  *dinfo = reinterpret_cast<int32_t*>(p);
  *counts = reinterpret_cast<int*>(p + 64);
  *lists = *counts + np;
  return 0;
}

int* lu_perm_area(int64_t n) {
  const int64_t np = (n + NBLK - 1) / NBLK;
  char* p = static_cast<char*>(ctx().ws[WS_PIVOT]);
  return reinterpret_cast<int*>(p + 64 + np * sizeof(int) + np * 4 * NBLK * sizeof(int));
}

// The steps of PLUDecBlock per panel, with a look-ahead of one panel: after panel k the main
// stream interchanges, solves (U12) and updates only the next panel's columns and factors that
// panel, while the auxiliary stream does the same for the columns after it and applies the
// interchanges left of the panel.  The panel is latency-bound (one exchange between workgroups
// per column), so the trailing DGEMM runs in its shadow.  Each element takes the same
// products as without look-ahead (the next panel's columns may go through another GEMM tile
// shape).  E0(k) = panel k done; E2(k) = the columns after the next panel updated, awaited
// before the main stream touches them:
//   main: panel(k); wait E2(k-1); swaps + U12 + update of the next panel's columns; panel(k+1)
//   aux:  wait E0(k); swaps on [0, j) and the far columns, their U12 and update; record E2(k)
int dgetrf_internal(int64_t n, double* a, int64_t lda, int32_t* dpiv, int32_t* info_host,
                    const LuScratch* sc) {
  hipStream_t s = stream();
  hipStream_t s2;
  hipEvent_t* ev;
  DDLA_RETURN_IF(aux_stream(&s2, &ev));
  int32_t* dinfo;
  int *counts, *lists;
  if (sc) {
    dinfo = sc->dinfo;
    counts = sc->counts;
    lists = sc->lists;
  } else {
    DDLA_RETURN_IF(lu_bookkeeping(n, &dinfo, &counts, &lists));
  }
  DDLA_RETURN_IF(launch_fill_u32(dinfo, 1, 0u, s));
  // the auxiliary stream starts after everything queued so far on the main stream
  DDLA_HIP_CHECK(hipEventRecord(ev[0], s));
  DDLA_HIP_CHECK(hipStreamWaitEvent(s2, ev[0], 0));
  bool pending = false;
  int k = 0;
  for (int64_t j = 0; j < n; j += NBLK, ++k) {
    const int b = (int)std::min<int64_t>(NBLK, n - j);
    const int64_t m = n - j;
    double* panel = a + j * lda + j;
    int* list_k = lists + (size_t)k * 4 * NBLK;
    PanelScratch ps{};
    if (sc) {  // a fresh tag range of the caller's area for each panel
      ps = PanelScratch{sc->panel_area, *sc->tag_counter + 1u};
      *sc->tag_counter += static_cast<unsigned>(b) + 2u;
    }
    DDLA_RETURN_IF(launch_getf2_panel_ex<double>(m, b, panel, lda, dpiv + j, dinfo, (int)j, list_k,
                                                 counts + k, s, sc ? &ps : nullptr));
    const int64_t rest = n - j - b;
    const int64_t nxt = std::min<int64_t>(NBLK, rest);  // the next panel's width
    const int64_t far0 = j + b + nxt, far = n - far0;   // columns after the next panel
    if (far > 0) DDLA_HIP_CHECK(hipEventRecord(ev[0], s));  // E0(k)
    if (pending) {
      DDLA_HIP_CHECK(hipStreamWaitEvent(s, ev[1], 0));  // E2(k-1): next panel's columns updated
      pending = false;
    }
    // interchanges on columns [0, j) and [j+b, n) of rows j..n-1 (LU.cs:316-329), here on the
    // next panel's columns (fused with its U12) or, for the last panels, on all of them
    if (far > 0)
      DDLA_RETURN_IF(launch_laswp_trsm_slab<double>(b, nxt, panel, lda, panel + b, lda, list_k,
                                                    counts + k, s));
    else
      DDLA_RETURN_IF(launch_laswp_list<double>(a + j * lda, lda, 0, n, j, j + b, list_k,
                                               counts + k, 2 * NBLK, s));
    if (rest > 0) {
      // A12 = trilu(A11) \ A12 (LU.cs:332)
      if (far <= 0)
        DDLA_RETURN_IF(launch_trsm_slab<double>(true, true, b, nxt, panel, lda, 1, panel + b, lda, s));
      // A22 -= A21 * A12 (LU.cs:334), the next panel's columns
      DDLA_RETURN_IF(launch_dgemm(rest, nxt, b, -1.0, panel + b * lda, lda, true, panel + b, lda,
                                  true, 1.0, panel + b * lda + b, lda, 0, s));
    }
    if (far > 0) {
      DDLA_HIP_CHECK(hipStreamWaitEvent(s2, ev[0], 0));
      DDLA_RETURN_IF(launch_laswp_list<double>(a + j * lda, lda, 0, n, j, far0, list_k, counts + k,
                                               2 * NBLK, s2));
      double* a12 = a + j * lda + far0;
      DDLA_RETURN_IF(launch_trsm_slab<double>(true, true, b, far, panel, lda, 1, a12, lda, s2));
      DDLA_RETURN_IF(launch_dgemm(rest, far, b, -1.0, panel + b * lda, lda, true, a12, lda, true,
                                  1.0, a12 + b * lda, lda, 0, s2));
      DDLA_HIP_CHECK(hipEventRecord(ev[1], s2));  // E2(k)
      pending = true;
    }
  }
  if (pending) DDLA_HIP_CHECK(hipStreamWaitEvent(s, ev[1], 0));
  DDLA_RETURN_IF(read_info(dinfo, info_host));
  PanelScratch ps{sc ? sc->panel_area : nullptr, 0};
  return check_abort(panel_abort_word(sc ? &ps : nullptr), s, "LU panel");
}

int dgetrs_internal(bool trans, int64_t n, int64_t nrhs, const double* a, int64_t lda,
                    const int32_t* dpiv, double* b, int64_t ldb) {
  hipStream_t s = stream();
  if (n <= 0 || nrhs <= 0) return 0;
  void* ws = nullptr;
  const size_t perm_bytes = ((size_t)n * sizeof(int) + 255) / 256 * 256;
  DDLA_RETURN_IF(workspace(WS_REFINE, perm_bytes + (size_t)n * nrhs * sizeof(double), &ws));
  int* perm = static_cast<int*>(ws);
  double* tmp = reinterpret_cast<double*>(static_cast<char*>(ws) + perm_bytes);
  DDLA_RETURN_IF(launch_piv_to_perm(n, dpiv, perm, s));
  if (!trans) {
    // X = U^-1 L^-1 P B
    DDLA_RETURN_IF(launch_gather_rows<double>(n, nrhs, b, ldb, perm, tmp, nrhs, false, s));
    DDLA_RETURN_IF(launch_dcopy2d(n, nrhs, tmp, nrhs, 1, b, ldb, s));
    DDLA_RETURN_IF(tri_solve(true, true, false, n, nrhs, a, lda, b, ldb));
    DDLA_RETURN_IF(tri_solve(false, false, false, n, nrhs, a, lda, b, ldb));
  } else {
    // A^T X = B: U^T V = B, L^T Y = V, X = P^-1 Y
    DDLA_RETURN_IF(tri_solve(false, false, true, n, nrhs, a, lda, b, ldb));
    DDLA_RETURN_IF(tri_solve(true, true, true, n, nrhs, a, lda, b, ldb));
    DDLA_RETURN_IF(launch_gather_rows<double>(n, nrhs, b, ldb, perm, tmp, nrhs, true, s));
    DDLA_RETURN_IF(launch_dcopy2d(n, nrhs, tmp, nrhs, 1, b, ldb, s));
  }
  return solves_done(nrhs, s);
}

// LUDecBlock (LU.cs:169-186): per 128-block, LU without pivoting of the diagonal block,
// A12 <- L11^-1 A12, A21 <- A21 U11^-1, A22 <- A22 - A21 A12.
int dgetrf_nopiv_internal(int64_t n, double* a, int64_t lda, int32_t* info_host) {
  hipStream_t s = stream();
  int32_t* dinfo;
  int *counts, *lists;
  DDLA_RETURN_IF(lu_bookkeeping(1, &dinfo, &counts, &lists));
  DDLA_HIP_CHECK(hipMemsetAsync(dinfo, 0, sizeof(int32_t), s));
  for (int64_t j = 0; j < n; j += NBLK) {
    const int b = (int)std::min<int64_t>(NBLK, n - j);
    double* a11 = a + j * lda + j;
    DDLA_RETURN_IF(launch_getf2_nopiv<double>(b, a11, lda, dinfo, (int)j, s));
    const int64_t rest = n - j - b;
    if (rest > 0) {
      DDLA_RETURN_IF(launch_trsm_slab<double>(true, true, b, rest, a11, lda, 1, a11 + b, lda, s));
      DDLA_RETURN_IF(launch_trsm_rlt_panel<double>(b, rest, a11, lda, a11 + b * lda, lda, s, true));
      DDLA_RETURN_IF(launch_dgemm(rest, rest, b, -1.0, a11 + b * lda, lda, true, a11 + b, lda, true,
                                  1.0, a11 + b * lda + b, lda, 0, s));
    }
  }
  return read_info(dinfo, info_host);
}

// LUSolveInternal (LU.cs:205-231)
int dgetrs_nopiv_internal(bool trans, int64_t n, int64_t nrhs, const double* a, int64_t lda,
                          double* b, int64_t ldb) {
  if (!trans) {
    DDLA_RETURN_IF(tri_solve(true, true, false, n, nrhs, a, lda, b, ldb));
    DDLA_RETURN_IF(tri_solve(false, false, false, n, nrhs, a, lda, b, ldb));
  } else {
    DDLA_RETURN_IF(tri_solve(false, false, true, n, nrhs, a, lda, b, ldb));
    DDLA_RETURN_IF(tri_solve(true, true, true, n, nrhs, a, lda, b, ldb));
  }
  return solves_done(nrhs, stream());
}

// Blocked lower Cholesky (CholeskyLowerBlock, Cholesky.cs:142-160) with a look-ahead of one
// block column: after L21 of step k, the main stream updates only the next block column and
// goes on to its diagonal block and L21, while the auxiliary stream updates the rest of the
// trailing triangle.  E0(k) = L21 of step k ready; E2(k) = far update of step k done, awaited
// before the main stream touches columns that update also covers.  Each element takes the
// same products as without look-ahead (the split only decides which launch computes it).
int dpotrf_internal(bool lower, int64_t n, double* a, int64_t lda, int32_t* info_host) {
  hipStream_t s = stream();
  hipStream_t s2;
  hipEvent_t* ev;
  DDLA_RETURN_IF(aux_stream(&s2, &ev));
  int32_t* dinfo;
  int *counts, *lists;
  DDLA_RETURN_IF(lu_bookkeeping(1, &dinfo, &counts, &lists));
  DDLA_HIP_CHECK(hipMemsetAsync(dinfo, 0, sizeof(int32_t), s));
  if (!lower) DDLA_RETURN_IF(launch_dtranspose_inplace(n, a, lda, s));
  DDLA_HIP_CHECK(hipEventRecord(ev[0], s));
  DDLA_HIP_CHECK(hipStreamWaitEvent(s2, ev[0], 0));
  bool pending = false;
  for (int64_t j = 0; j < n; j += NBLK) {
    const int b = (int)std::min<int64_t>(NBLK, n - j);
    double* a11 = a + j * lda + j;
    DDLA_RETURN_IF(launch_potf2_ex<double>(b, a11, lda, dinfo, (int)j, s));
    const int64_t rest = n - j - b;
    if (rest <= 0) break;
    double* a21 = a11 + b * lda;
    // A21 = A21 * L11^-T (Cholesky.cs:156-157)
    DDLA_RETURN_IF(launch_trsm_rlt_panel<double>(b, rest, a11, lda, a21, lda, s, false));
    // lower(A22) -= A21 * A21^T (Cholesky.cs:158): the next block column here, the rest on
    // the auxiliary stream
    const int64_t nxt = std::min<int64_t>(NBLK, rest), far = rest - nxt;
    if (far > 0) DDLA_HIP_CHECK(hipEventRecord(ev[0], s));  // E0(k)
    if (pending) {
      DDLA_HIP_CHECK(hipStreamWaitEvent(s, ev[1], 0));  // E2(k-1)
      pending = false;
    }
    DDLA_RETURN_IF(launch_dgemm(rest, nxt, b, -1.0, a21, lda, true, a21, lda, false, 1.0, a21 + b,
                                lda, 1, s));
    if (far > 0) {
      DDLA_HIP_CHECK(hipStreamWaitEvent(s2, ev[0], 0));
      double* af = a21 + nxt * lda;
      DDLA_RETURN_IF(launch_dgemm(far, far, b, -1.0, af, lda, true, af, lda, false, 1.0,
                                  af + b + nxt, lda, 1, s2));
      DDLA_HIP_CHECK(hipEventRecord(ev[1], s2));  // E2(k)
      pending = true;
    }
  }
  if (pending) DDLA_HIP_CHECK(hipStreamWaitEvent(s, ev[1], 0));
  if (!lower) DDLA_RETURN_IF(launch_dtranspose_inplace(n, a, lda, s));
  return read_info(dinfo, info_host);
}

int dpotrs_internal(bool lower, int64_t n, int64_t nrhs, const double* a, int64_t lda, double* b,
                    int64_t ldb) {
  if (lower) {
    DDLA_RETURN_IF(tri_solve(true, false, false, n, nrhs, a, lda, b, ldb));
    DDLA_RETURN_IF(tri_solve(true, false, true, n, nrhs, a, lda, b, ldb));
  } else {
    DDLA_RETURN_IF(tri_solve(false, false, true, n, nrhs, a, lda, b, ldb));
    DDLA_RETURN_IF(tri_solve(false, false, false, n, nrhs, a, lda, b, ldb));
  }
  return solves_done(nrhs, stream());
}

// Per 128-column block: the diagonal block's LDL^T (one workgroup), A21 <- A21 L11^-T (unit
// diagonal), Y21 = A21 and A21 <- A21 D11^-1 (the multipliers), then the lower triangle of A22
// takes -Y21 A21^T on the GEMM (GeMMT, LDLT.cs:185-193).  A zero pivot d_k gives info k + 1.
int dldlt_internal(bool lower, int64_t n, double* a, int64_t lda, int32_t* info_host) {
  hipStream_t s = stream();
  int32_t* dinfo;
  int *counts, *lists;
  DDLA_RETURN_IF(lu_bookkeeping(1, &dinfo, &counts, &lists));
  DDLA_RETURN_IF(launch_fill_u32(dinfo, 1, 0u, s));
  double* y = nullptr;
  if (n > NBLK) {
    void* ws = nullptr;
    DDLA_RETURN_IF(workspace(WS_TEMP, (size_t)(n - NBLK) * NBLK * sizeof(double), &ws));
    y = static_cast<double*>(ws);
  }
  if (!lower) DDLA_RETURN_IF(launch_dtranspose_inplace(n, a, lda, s));
  for (int64_t j = 0; j < n; j += NBLK) {
    const int b = (int)std::min<int64_t>(NBLK, n - j);
    double* a11 = a + j * lda + j;
    DDLA_RETURN_IF(launch_ldlt_diag(b, a11, lda, dinfo, (int)j, s));
    const int64_t rest = n - j - b;
    if (rest <= 0) break;
    double* a21 = a11 + b * lda;
    DDLA_RETURN_IF(launch_trsm_rlt_panel<double>(b, rest, a11, lda, a21, lda, s, false, true));
    DDLA_RETURN_IF(launch_ldlt_scale(rest, b, a11, lda, a21, y, b, s));
    // lower(A22) -= Y21 * A21^T
    DDLA_RETURN_IF(launch_dgemm(rest, rest, b, -1.0, y, b, true, a21, lda, false, 1.0, a21 + b,
                                lda, 1, s));
  }
  if (!lower) DDLA_RETURN_IF(launch_dtranspose_inplace(n, a, lda, s));
  return read_info(dinfo, info_host);
}

// lower: X <- L^-1 X, X <- D^-1 X, X <- L^-T X; upper: the same with L = U^T
int dldlt_solve_internal(bool lower, int64_t n, int64_t nrhs, const double* a, int64_t lda,
                         double* b, int64_t ldb) {
  DDLA_RETURN_IF(tri_solve(lower, true, !lower, n, nrhs, a, lda, b, ldb));
  DDLA_RETURN_IF(launch_scale_rows_by_inv_diag(n, nrhs, a, lda, b, ldb, stream()));
  DDLA_RETURN_IF(tri_solve(lower, true, lower, n, nrhs, a, lda, b, ldb));
  return solves_done(nrhs, stream());
}

}  // namespace ddla
