// Internal declarations shared by the HIP translation units of libddla_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/ddla_native.h"

namespace ddla {

// ------------------------------------------------------------------ errors ---------------
void set_error(const char* fmt, ...);
const char* last_error();

#define DDLA_HIP_CHECK(expr)                                                      \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      ::ddla::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),    \
                        __FILE__, __LINE__);                                      \
      return DDLA_ERR_HIP;                                                        \
    }                                                                             \
  } while (0)

#define DDLA_ARG_CHECK(cond, ...)                 \
  do {                                            \
    if (!(cond)) {                                \
      ::ddla::set_error(__VA_ARGS__);             \
      return DDLA_ERR_INVALID_ARG;                \
    }                                             \
  } while (0)

#define DDLA_RETURN_IF(st)       \
  do {                           \
    int _s = (st);               \
    if (_s != 0) return _s;      \
  } while (0)

// ---------------------------------------------------------------- context ----------------
constexpr unsigned DEFAULT_WAIT_LIMIT = 1u << 24;

enum WorkspaceSlot {
  WS_PANEL = 0,   // panel-kernel exchange slots
  WS_PIVOT = 1,   // pivots / permutations / info words
  WS_TEMP = 2,    // general temporaries (packed operands, transposes)
  WS_MIXED = 3,   // low-precision working copies
  WS_MIXED2 = 4,
  WS_REFINE = 5,  // refinement vectors
  WS_GMRES = 6,
  WS_SMALL = 7,   // small reduction outputs
  WS_DINV = 8,    // inverted diagonal blocks for the FP64 solves
  WS_SPLIT = 9,   // low-precision parts of the trailing-update operands
  WS_SPLIT2 = 10, // ... for updates on the auxiliary stream
  WS_TEMP2 = 11,  // row-interchange staging on the auxiliary stream
  WS_TSLU = 12,   // tournament-pivoting tree (candidate sets, arrival words)
  WS_SPLITP = 13, // split panel operand shared by the updates of one panel (two in flight)
  WS_BLAS3 = 14,  // level-3 BLAS: the expanded triangular / symmetric operand
  WS_BLAS3B = 15, // level-3 BLAS: the product before it is written back over B
  WS_QR_SYNC = 16, // QR panel exchange area
  WS_QR_WORK = 17, // QR: explicit Householder vectors of a block, products, tau
  WS_QR_Y = 18,    // QR solve: Q^T B
  WS_SPLITK = 19,  // split-K GEMM partial products (main stream only)
  WS_EIN_IDX = 20, // tensor contraction: scatter vectors
  WS_EIN_A = 21,   // tensor contraction: gathered operands and product
  WS_EIN_B = 22,
  WS_EIN_C = 23,
  WS_PIVLIST = 24, // relative pivots as per-block row-move lists (launch_piv_to_perm)
  WS_COUNT = 25,
};

struct Context {
  bool ready = false;
  int device = 0;
  int num_cus = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  // second stream of the library's own (look-ahead: trailing updates overlap the next panel),
  // default priority like the main one; events order the two
  hipStream_t aux_stream = nullptr;
  hipEvent_t ev[4] = {};
  // small pinned host buffer (asynchronous device-to-host copies of status words)
  void* pinned = nullptr;
  // growable scratch buffers (device memory kept for the life of the context)
  void* ws[WS_COUNT] = {};
  size_t ws_bytes[WS_COUNT] = {};
  // bytes of each slot cleared since it was allocated, and the tag sequence of the exchange
  // words kept there (kernels that wait on each other's words compare 32-bit tags; a cleared
  // slot holds no live tag)
  size_t ws_cleared[WS_COUNT] = {};
  unsigned ws_tag[WS_COUNT] = {};
  // poll rounds a cross-workgroup wait (LU panel, triangular solve) makes before it gives up
  unsigned wait_limit = DEFAULT_WAIT_LIMIT;
  // set once a caller changed wait_limit: from then on every panel launch first writes the
  // limit into its exchange area, where the kernel reads it (a cleared word means the default)
  bool wait_limit_set = false;
};

Context& ctx();
int ensure_init();
// Locates (hipGetSymbolAddress) and clears the static scratch of the no-allocation entry
// points on `device`, which must be current; called once per device by the library's set-up.
int noalloc_setup(int device, hipStream_t s);
hipStream_t stream();
// the auxiliary stream and its events (created on first use)
int aux_stream(hipStream_t* s, hipEvent_t** ev);
// the context's pinned host buffer (allocated on first use, kept for the process)
constexpr size_t PINNED_BYTES = 4096;
int pinned_host(void** p);
// Returns a device scratch buffer `slot` of at least `bytes` (contents undefined).
int workspace(int slot, size_t bytes, void** out);
// The same, with the first `bytes` bytes cleared at least once since the buffer was
// allocated (on stream s, only when needed); *fresh tells the caller whether this call
// cleared it, in which case ctx().ws_tag[slot] has been reset to 0.
int workspace_cleared(int slot, size_t bytes, void** out, hipStream_t s, bool* fresh = nullptr);
// Next tag of slot's sequence (never 0); a wrapped sequence clears the slot again first.
int next_tag(int slot, size_t bytes, unsigned step, hipStream_t s, unsigned* tag);
// Kernels whose workgroups wait on each other's words (the LU panel, the persistent triangular
// solve) bound every wait by ctx().wait_limit poll rounds: a workgroup that reaches the bound
// stores its launch's tag in the launch's abort word, and every workgroup of the launch leaves
// at its next poll.  check_abort synchronises s, reads such a word and, when it was set, clears
// it and returns DDLA_ERR_HIP naming `what`.
int check_abort(unsigned* word, hipStream_t s, const char* what);


// Optional per-kernel timing (ddla_profile_*): while enabled, a ProfScope around a launch
// records HIP events on the launch stream and accumulates the kernel's time together with
// its algorithmic flops and bytes under `name`.
bool profiling();
class ProfScope {
 public:
  ProfScope(const char* name, double flops, double bytes, hipStream_t s);
  ~ProfScope();
  ProfScope(const ProfScope&) = delete;
  ProfScope& operator=(const ProfScope&) = delete;

 private:
  hipStream_t s_;
  int entry_ = -1;
  double flops_ = 0.0, bytes_ = 0.0;
  void* a_ = nullptr;
  void* b_ = nullptr;
};
void profile_enable(bool on);
void profile_reset();
void profile_only(const char* name);
int profile_read(int idx, char* name, size_t len, int64_t* count, double* ms, double* flops,
                 double* bytes);

inline int ceil_div(int64_t a, int64_t b) { return static_cast<int>((a + b - 1) / b); }

// Maps a linear workgroup id to a (tile_m, tile_n) pair of a GEMM: workgroup ids are
// dispatched round-robin over the 8 XCDs, so ids are first regrouped into XCD-contiguous
// bands; each band walks a grouped order (8 tile rows at a time) so the A row panels and
// B column panels it reuses stay in that XCD's L2.
__device__ __forceinline__ void gemm_tile_of(int pid, int tiles_m, int tiles_n, int& tm, int& tn) {
  const int T = tiles_m * tiles_n;
  const int T8 = (T / 8) * 8;
  if (pid < T8) pid = (pid % 8) * (T8 / 8) + pid / 8;
  constexpr int G = 8;
  const int per_group = G * tiles_n;
  const int group = pid / per_group;
  const int first_m = group * G;
  const int gm = min(tiles_m - first_m, G);
  const int in_group = pid % per_group;
  tm = first_m + in_group % gm;
  tn = in_group / gm;
}

// Tiles (i, j) with j <= i (lower triangle of C, square tiles, C's diagonal through tile (0, 0))
// of a tiles_m x tiles_n grid: rows i < D = min(tiles_m, tiles_n) hold i + 1 of them, later
// rows all tiles_n.
__host__ __device__ inline int64_t lower_tile_count(int tiles_m, int tiles_n) {
  const int64_t d = tiles_m < tiles_n ? tiles_m : tiles_n;
  return d * (d + 1) / 2 + (int64_t)(tiles_m - d) * tiles_n;
}
// The pid-th of those tiles, spread over the XCDs as gemm_tile_of spreads a full grid: XCD x
// (pid % 8 == x) takes a contiguous range of the row-major enumeration, so every XCD gets the
// same number of tiles (a full-grid order with the upper tiles skipped leaves the XCD of the
// last rows with most of the work).
__device__ __forceinline__ void lower_tile_of(int pid, int tiles_m, int tiles_n, int& tm,
                                              int& tn) {
  const int T = (int)lower_tile_count(tiles_m, tiles_n);
  const int T8 = (T / 8) * 8;
  const int t = pid < T8 ? (pid % 8) * (T8 / 8) + pid / 8 : pid;
  const int d = tiles_m < tiles_n ? tiles_m : tiles_n;
  const int t1 = d * (d + 1) / 2;
  if (t < t1) {
    int i = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    while (i * (i + 1) / 2 > t) --i;
    tm = i;
    tn = t - i * (i + 1) / 2;
  } else {
    const int u = t - t1;
    tm = d + u / tiles_n;
    tn = u % tiles_n;
  }
}

// ------------------------------------------------------------ kernels (launchers) --------
// FP64 GEMM on row-major C: C = alpha*A*B + beta*C, A m x k, B k x n.
//   a_kcontig: A(i,k) at A[i*lda + k], else A[i + k*lda]
//   b_ncontig: B(k,j) at B[k*ldb + j], else B[k + j*ldb]
//   tri: 0 full, 1 lower triangle of C only (i >= j), 2 upper (i <= j)
int launch_dgemm(int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                 bool a_kcontig, const double* B, int64_t ldb, bool b_ncontig, double beta,
                 double* C, int64_t ldc, int tri, hipStream_t s);

int launch_sgemm(int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                 bool a_kcontig, const float* B, int64_t ldb, bool b_ncontig, float beta, float* C,
                 int64_t ldc, int tri, hipStream_t s);

// C += alpha*A*B on the FP16 (bf16=false) or BF16 matrix cores, FP32 C, FP32 operands
// rounded per 128-block with power-of-two scales; scale_ws holds >= tiles_m + tiles_n floats
int launch_hgemm(bool bf16, int terms, bool b_kcontig, int tri, int64_t m, int64_t n, int64_t k,
                 float alpha, const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                 int64_t ldc, float* scale_ws, hipStream_t s, int ws_slot = WS_SPLIT);

// the same in two parts, for an operand shared by several updates: split once, multiply many
// times.  Low-precision parts of an operand: element (r, k) of part t at
// parts[t plane + r Kp + k] (Kp = k rounded up to 32), one scale per 128 rows.
struct SplitOp {
  const void* parts;
  int64_t plane;
  const float* scales;
};
int launch_split(bool bf16, int terms, bool kcontig, int64_t rows, int64_t k, const float* src,
                 int64_t ld, void* out, float* scales, hipStream_t s);
int launch_hgemm_parts(bool bf16, int terms, int tri, int64_t m, int64_t n, int64_t k,
                       float alpha, SplitOp a, SplitOp b, float* C, int64_t ldc, hipStream_t s);

// ------------------------------------------------------------ host-side helpers (api.hip) --
bool valid_trans(int t);
// real data: conjugation is the identity (BLIS ignores conj for real types)
bool is_trans(int t);
void pack_host(int64_t m, int64_t n, const double* src, int64_t rs, int64_t cs, double* dst);
void unpack_host(int64_t m, int64_t n, const double* src, double* dst, int64_t rs, int64_t cs);
// device buffer owned by a host-pointer entry point for the duration of the call
struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  int alloc(size_t bytes);
  double* d() const { return static_cast<double*>(p); }
};
// strided host view (m x n, rs, cs) -> new contiguous row-major device buffer, and back
int upload_view(int64_t m, int64_t n, const double* host, int64_t rs, int64_t cs, DevBuf& buf);
int download_view(int64_t m, int64_t n, const DevBuf& buf, double* host, int64_t rs, int64_t cs);
// C = alpha op(A) op(B) + beta C on device pointers with any strides (BLIS conventions);
// tri (m == n): 0 all of C, 1 / 2 only its lower / upper triangle
int gemm_strided(int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
                 const double* a, int64_t rsa, int64_t csa, const double* b, int64_t rsb,
                 int64_t csb, double beta, double* c, int64_t rsc, int64_t csc, int tri);
// B = alpha op(A)^-1 B (left) or alpha B op(A)^-1 (right), row-major device matrices
int trsm_rowmajor(int side, int uplo, int transa, int diag, int64_t m, int64_t n, double alpha,
                  const double* a, int64_t lda, double* b, int64_t ldb);

// general strided copy / transpose helpers
int launch_dcopy2d(int64_t m, int64_t n, const double* src, int64_t rs, int64_t cs, double* dst,
                   int64_t ld, hipStream_t s);
// dst(i,j) at dst[i*rs + j*cs] = src[i*ld + j]
int launch_dscatter2d(int64_t m, int64_t n, const double* src, int64_t ld, double* dst, int64_t rs,
                      int64_t cs, hipStream_t s);
int launch_dscale2d(int64_t m, int64_t n, double beta, double* c, int64_t rs, int64_t cs,
                    int tri, hipStream_t s);

// p[0 .. count) = v (a kernel, not hipMemsetAsync, so that a captured graph replays it)
int launch_fill_u32(void* p, int64_t count, unsigned v, hipStream_t s);

// LU pieces (lu_kernels.hip).  The panel kernel writes relative pivots to piv[0..nb),
// sets *info = info_base + i + 1 at the first zero pivot (if *info was 0) and emits the
// list of moved rows as (dst, src) pairs relative to the panel's first row.
// ext: a caller-provided exchange area of panel_area_bytes() (all zero, or holding only words
// with tags below first_tag) and the first tag of the launch's range first_tag .. first_tag +
// nb + 1; null: the library's workspace and tag sequence.
struct PanelScratch {
  void* area;
  unsigned first_tag;
};
size_t panel_area_bytes();
// the abort word of the panel launches in ext's area (null: the library's workspace)
unsigned* panel_abort_word(const PanelScratch* ext);
template <typename T>
int launch_getf2_panel_ex(int64_t m, int nb, T* P, int64_t ld, int32_t* piv, int32_t* info,
                          int info_base, int* list_out, int* count_out, hipStream_t s,
                          const PanelScratch* ext = nullptr);
// applies such a list (at most 2 * 128 entries) to columns [0, ncols) minus
// [skip_lo, skip_hi) of the rows from row0 on
template <typename T>
int launch_laswp_list(T* A, int64_t ld, int64_t row0, int64_t ncols, int64_t skip_lo,
                      int64_t skip_hi, const int* list, const int* count, int max_count,
                      hipStream_t s);
int launch_perm_compose(int n, int npanels, int nb, const int* lists, const int* counts, int* perm,
                        hipStream_t s);
// Tournament-pivoting panel (tslu.hip, FP32): same outputs as the partial-pivoting panel
// (factored panel, *info, (dst, src) list) except relative pivots; nb must be 128.
bool tslu_supported(int64_t m, int nb);
int launch_tslu_panel(int64_t m, int nb, float* P, int64_t ld, int32_t* info, int info_base,
                      int* list_out, int* count_out, hipStream_t s);

// triangular block solves (trsm.hip): op(T) X = B for one bs x bs triangle (bs <= 128)
template <typename T>
int launch_laswp_trsm_slab(int bs, int64_t ncols, const T* L, int64_t ldl, T* B, int64_t ldb,
                           const int* list, const int* count, hipStream_t s);
template <typename T>
int launch_trsm_slab(bool lower, bool unit, int bs, int64_t ncols, const T* A, int64_t rs,
                     int64_t cs, T* B, int64_t ldb, hipStream_t s);
// op(A) X = B, A n x n row-major; lower/upper name the stored triangle of A
template <typename T>
int trsm_left(bool lower, bool unit, bool transA, int64_t n, int64_t nrhs, const T* A,
              int64_t lda, T* B, int64_t ldb, hipStream_t s);

// persistent triangular solves with <= 8 right-hand sides (trsv.hip); lower/unit/trans
// describe op(T) for trsv_prepare and (lower of op(T), trans) for trsv_solve
size_t trsv_dinv_elems(int64_t n);
template <typename T>
int trsv_prepare(bool lower, bool unit, bool trans, int64_t n, const T* A, int64_t lda, T* Dinv,
                 hipStream_t s);
// skip (device word, may be null): when nonzero at the start of the launch the solve does nothing
template <typename T>
int trsv_solve(bool lower, bool trans, int64_t n, int nrhs, const T* A, int64_t lda, const T* Dinv,
               T* B, int64_t ldb, hipStream_t s, const int* skip = nullptr);
// the abort word of the persistent triangular solves (null before the first solve)
unsigned* trsv_abort_word();
// the same solve for many right-hand sides (FP32): GEMMs on the prepared diagonal inverses
int trsv_solve_blocked(bool lower, bool trans, int64_t n, int nrhs, const float* A, int64_t lda,
                       const float* Dinv, float* B, int64_t ldb, hipStream_t s);

// Cholesky pieces (chol_kernels.hip)
template <typename T>
int launch_potf2_ex(int nb, T* A, int64_t lda, int32_t* info, int info_base, hipStream_t s);
// B <- B L^-T for lower L (l_transposed: B <- B U^-1 for the upper U stored at L); unit: the
// triangle's diagonal is taken as ones (not read)
template <typename T>
int launch_trsm_rlt_panel(int nb, int64_t nrows, const T* L, int64_t ldl, T* B, int64_t ldb,
                          hipStream_t s, bool l_transposed, bool unit = false);
// LU without pivoting of one nb x nb diagonal block (nb <= 128)
template <typename T>
int launch_getf2_nopiv(int nb, T* A, int64_t lda, int32_t* info, int info_base, hipStream_t s);
int launch_dtranspose_inplace(int64_t n, double* a, int64_t lda, hipStream_t s);
// LDL^T pieces (chol_kernels.hip): the nb x nb diagonal block (nb <= 128, lower triangle);
// Y = A21, A21 *= diag(A11)^-1 by columns; X[i, :] *= 1 / A[i, i]
int launch_ldlt_diag(int nb, double* A, int64_t lda, int32_t* info, int info_base,
                     hipStream_t s);
int launch_ldlt_scale(int64_t rows, int b, const double* A11, int64_t lda, double* A21,
                      double* Y, int64_t ldy, hipStream_t s);
int launch_scale_rows_by_inv_diag(int64_t n, int64_t nrhs, const double* A, int64_t lda,
                                  double* X, int64_t ldx, hipStream_t s);

// relative pivots that stay inside n rows (0 <= i + piv[i] < n wherever piv[i] > 0), checked on
// the device; synchronises s and returns DDLA_ERR_INVALID_ARG otherwise
int check_pivots(int64_t n, const int32_t* piv, hipStream_t s);
// pivots -> permutation (perm[i] = source row of row i after all interchanges)
int launch_piv_to_perm(int64_t n, const int32_t* piv, int32_t* perm, hipStream_t s);
template <typename T>
int launch_gather_rows(int64_t n, int64_t ncols, const T* src, int64_t lds,
                       const int32_t* perm, T* dst, int64_t ldd, bool inverse, hipStream_t s);

// internal FP64 drivers
// Device memory a factorization may take from its caller instead of the library's workspaces
// (the no-allocation entry point): the panel exchange area, the info word, per-panel counts
// (npanels ints) and row lists (npanels * 4 * 128 ints), and the tag counter of the area.
struct LuScratch {
  void* panel_area;
  int32_t* dinfo;
  int* counts;
  int* lists;
  unsigned* tag_counter;  // host word: last tag used in panel_area
};
int dgetrf_internal(int64_t n, double* a, int64_t lda, int32_t* dpiv, int32_t* info_host,
                    const LuScratch* sc = nullptr);
int dgetrs_internal(bool trans, int64_t n, int64_t nrhs, const double* a, int64_t lda,
                    const int32_t* dpiv, double* b, int64_t ldb);
int dpotrf_internal(bool lower, int64_t n, double* a, int64_t lda, int32_t* info_host);
int dldlt_internal(bool lower, int64_t n, double* a, int64_t lda, int32_t* info_host);
// Householder QR in the UT form (qr.hip): T has qr_t_rows(m, n) rows and n columns
int qr_t_rows(int64_t m, int64_t n);
int dgeqrf_internal(int64_t m, int64_t n, double* a, int64_t lda, double* t, int64_t ldt);
int dqr_apply_internal(bool trans, int64_t m, int64_t n, const double* a, int64_t lda,
                       const double* t, int64_t ldt, int64_t nrhs, double* b, int64_t ldb);
int dcontract_internal(const char* expr, double alpha, const double* a, int na,
                       const int64_t* len_a, const int64_t* str_a, const double* b, int nb,
                       const int64_t* len_b, const int64_t* str_b, double beta, double* c,
                       int nc, const int64_t* len_c, const int64_t* str_c);
int dqr_solve_internal(int64_t m, int64_t n, int64_t nrhs, const double* a, int64_t lda,
                       const double* t, int64_t ldt, const double* b, int64_t ldb, double* x,
                       int64_t ldx);
int dldlt_solve_internal(bool lower, int64_t n, int64_t nrhs, const double* a, int64_t lda,
                         double* b, int64_t ldb);
int dpotrs_internal(bool lower, int64_t n, int64_t nrhs, const double* a, int64_t lda, double* b,
                    int64_t ldb);
int dgetrf_nopiv_internal(int64_t n, double* a, int64_t lda, int32_t* info_host);
int dgetrs_nopiv_internal(bool trans, int64_t n, int64_t nrhs, const double* a, int64_t lda,
                          double* b, int64_t ldb);

// batched uneven factorizations (batched.hip)
struct BatchPlan;
int batch_create(BatchPlan** out, int prec, int64_t count, const int32_t* n_host,
                 const int32_t* kind_host, void* a, int32_t* piv, int32_t* info, int tune);
int batch_execute(BatchPlan* p);
int batch_destroy(BatchPlan* p);
int batch_describe(const BatchPlan* p, char* buf, size_t len);

}  // namespace ddla
