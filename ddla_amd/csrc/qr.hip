// Blocked Householder QR in the UT form (QR.QRDecompose / ApplyQlhfc / ApplyQlnfc / Solve,
// DDLA/Factorizations/QR.cs:130-149, :326-521; Householder vectors as in
// DDLA/Transformations/HouseHolder.cs:34-111).
//
// Layout: A row-major (m x n, lda); T row-major with qr_t_rows(m, n) = min(256, max(m, n))
// rows and n columns; block k (columns k .. k + b - 1) keeps its b x b upper-triangular factor
// in T[0:b, k:k+b]: tau_j on the diagonal, u_i^T u_j above it (the reference's layout).
//
// Factorization, per 256-column block:
//   * the panel in sub-panels of 32 columns, each by qr_panel_kernel: one workgroup of 256
//     threads per 256 rows, one row per thread with its 32 entries in registers.  Per column
//     the workgroups exchange (through tagged words, as the LU panel does) their partial sums
//     of squares, then their partial products w = a12 + A22^T u2; every workgroup sums the
//     partials in the same fixed order, so all of them compute bit-identical Householder
//     scalars;
//   * the sub-panel's block reflector applied to the rest of the block on the GEMM
//     (W = U^T C, W = T^-T W, C -= U W), with its 32 x 32 T from U^T U on the GEMM;
//   * the block's full T factor (U^T U on the GEMM, strictly upper part, tau on the diagonal)
//     and the block reflector applied to the trailing columns the same way.
#include <algorithm>

#include "ddla_internal.hpp"

namespace ddla {
namespace {

typedef unsigned long long u64;

constexpr int QR_NB = 256;     // QR.BlockSize (QR.cs:16)
constexpr int SW = 32;         // sub-panel width
constexpr int ROWS = 256;      // rows per workgroup (one per thread)
constexpr int WMAX = 128;      // workgroups of one panel launch: m <= 32768
constexpr int ASLOT = 16;      // words per phase-A slot (one 128-byte line)

// exchange area (64-bit words): phase A [WMAX][ASLOT] (sum of squares hi/lo, chi hi/lo),
// phase B [WMAX][SW][2] (partial products), then the abort pair (word 0: first tag of a launch
// that gave up; word 1: poll-round limit, 0 meaning DEFAULT_WAIT_LIMIT)
constexpr size_t QA_WORDS = (size_t)WMAX * ASLOT;
constexpr size_t QB_WORDS = (size_t)WMAX * SW * 2;
constexpr size_t QR_ABORT_OFF = (QA_WORDS + QB_WORDS) * sizeof(u64);
constexpr size_t QR_AREA_BYTES = QR_ABORT_OFF + 256;

__device__ __forceinline__ void put(u64* p, unsigned payload, unsigned tag) {
  __hip_atomic_store(p, (static_cast<u64>(payload) << 32) | tag, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 get(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void put_d(u64* p, double v, unsigned tag) {
  const u64 b = __builtin_bit_cast(u64, v);
  put(p, static_cast<unsigned>(b >> 32), tag);
  put(p + 1, static_cast<unsigned>(b), tag);
}
__device__ __forceinline__ double join_d(u64 hi, u64 lo) {
  return __builtin_bit_cast(double, ((hi >> 32) << 32) | (lo >> 32));
}
__device__ __forceinline__ bool tagged(u64 w, unsigned tag) {
  return static_cast<unsigned>(w) == tag;
}

// Every 32nd poll round of a thread reads the abort pair: past the limit it stores `id` (the
// launch's first tag, never 0) and gives up; finding `id` there it gives up too.
__device__ __forceinline__ bool wait_expired(u64* area, unsigned id, unsigned rounds) {
  u64* pair = area + QR_ABORT_OFF / sizeof(u64);
  const u64 w = get(pair);
  if (static_cast<unsigned>(w) == id) return true;
  const unsigned lim = static_cast<unsigned>(w >> 32);
  if (rounds < (lim ? lim : DEFAULT_WAIT_LIMIT)) return false;
  __hip_atomic_store(reinterpret_cast<unsigned*>(pair), id, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  return true;
}
__device__ __forceinline__ bool keep_waiting(u64* area, unsigned id, unsigned& rounds) {
  ++rounds;
  return (rounds & 31u) != 0 || !wait_expired(area, id, rounds);
}

// butterfly sum over the 64 lanes: every lane ends with the same value (each level adds two
// equal-order pairs, so the association is the same in every lane)
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// Householder QR of the mrows x sw sub-panel at A (sw <= 32, row-major, lda), in place:
// R on and above the diagonal, u2 below, tau_j -> taus[j].  gridDim.x = ceil(mrows / 256)
// workgroups, all resident together (at most WMAX).
__global__ __launch_bounds__(ROWS) void qr_panel_kernel(int mrows, int sw, double* __restrict__ A,
                                                        int64_t lda, double* __restrict__ taus,
                                                        u64* __restrict__ area, unsigned base) {
  __shared__ double wsum[ROWS / 64][SW];
  __shared__ double part[8][SW];
  __shared__ double wvec[SW];
  __shared__ double sc[4];  // chi of the column, total sum of squares
  __shared__ int sfail;
  const int W = gridDim.x, w = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row = w * ROWS + tid;
  const bool own = row < mrows;
  u64* qa = area;
  u64* qb = area + QA_WORDS;
  const unsigned id = base + 1u;

  double v[SW];
#pragma unroll
  for (int c = 0; c < SW; ++c) v[c] = (own && c < sw) ? A[(int64_t)row * lda + c] : 0.0;
  if (tid == 0) sfail = 0;
  unsigned rounds = 0;

  for (int jj = 0; jj < sw; ++jj) {
    const unsigned tag = base + static_cast<unsigned>(jj) + 1u;
    // ---- phase A: sum of squares of the column below row jj, and chi = a_jj
    double xj = 0.0;
#pragma unroll
    for (int c = 0; c < SW; ++c)
      if (c == jj) xj = v[c];
    const double sq = (own && row > jj) ? xj * xj : 0.0;
    const double ws = wave_sum(sq);
    if (lane == 0) wsum[wave][0] = ws;
    if (row == jj) sc[0] = xj;  // workgroup 0 only (jj < 32)
    __syncthreads();
    if (tid == 0) {
      u64* slot = qa + (size_t)w * ASLOT;
      put_d(slot, ((wsum[0][0] + wsum[1][0]) + wsum[2][0]) + wsum[3][0], tag);
      if (w == 0) put_d(slot + 2, sc[0], tag);
    }
    if (wave == 0) {
      bool failed = false;
      u64 a0, a1, b0, b1, c0, c1;
      const int s1 = lane + 64;
      for (;;) {
        const u64* p0 = qa + (size_t)min(lane, W - 1) * ASLOT;
        const u64* p1 = qa + (size_t)min(s1, W - 1) * ASLOT;
        a0 = get(p0);
        a1 = get(p0 + 1);
        b0 = get(p1);
        b1 = get(p1 + 1);
        c0 = get(qa + 2);
        c1 = get(qa + 3);
        const bool ok = tagged(a0, tag) && tagged(a1, tag) && tagged(b0, tag) &&
                        tagged(b1, tag) && tagged(c0, tag) && tagged(c1, tag);
        if (__all(ok)) break;
        if (!keep_waiting(area, id, rounds)) {
          failed = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      if (failed) {
        if (lane == 0) sfail = 1;
      } else {
        const double s = (lane < W ? join_d(a0, a1) : 0.0) + (s1 < W ? join_d(b0, b1) : 0.0);
        const double tot = wave_sum(s);
        if (lane == 0) {
          sc[1] = tot;
          sc[2] = join_d(c0, c1);
        }
      }
    }
    __syncthreads();
    if (sfail) break;
    // ---- the reflector (BuildHouseHolder), identical in every workgroup
    const double chi = sc[2];
    const double len2 = sqrt(sc[1]);
    double tau, rmiu = 0.0, newchi;
    if (len2 == 0.0) {
      newchi = -chi;
      tau = 0.5;
    } else {
      const double rho = -copysign(hypot(chi, len2), chi);
      const double miu = chi - rho;
      newchi = rho;
      rmiu = 1.0 / miu;
      const double s = len2 / fabs(miu);
      tau = (1.0 + s * s) / 2.0;
    }
    const double u = (own && row > jj) ? (len2 == 0.0 ? xj : xj * rmiu) : 0.0;
    const double coef = (row == jj) ? 1.0 : u;  // row jj: the implicit 1 of u
#pragma unroll
    for (int c = 0; c < SW; ++c)
      if (c == jj) v[c] = (row == jj) ? newchi : ((own && row > jj) ? u : v[c]);
    // ---- phase B: w_c = (a_jc + sum_{r > jj} a_rc u_r) / tau for jj < c < sw
    const int ncol = sw - jj - 1;
    if (ncol > 0) {
#pragma unroll
      for (int c = 0; c < SW; ++c) {
        if (c > jj && c < sw) {
          const double t = wave_sum(coef * v[c]);
          if (lane == 0) wsum[wave][c] = t;
        }
      }
      __syncthreads();
      if (tid > jj && tid < sw)
        put_d(qb + ((size_t)w * SW + tid) * 2,
              ((wsum[0][tid] + wsum[1][tid]) + wsum[2][tid]) + wsum[3][tid], tag);
      // thread: column jj + 1 + (tid & 31), slots g, g + 8, ... (g = tid >> 5), in order
      const int ci = tid & 31, g = tid >> 5;
      const int c = jj + 1 + ci;
      const bool hc = ci < ncol;
      double acc = 0.0;
      bool failed = false;
      for (int s0 = g; s0 < W; s0 += 8 * 4) {  // four slots per round trip
        u64 h[4], l[4];
        for (;;) {
          bool ok = true;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int s = s0 + 8 * q;
            const u64* p = qb + ((size_t)min(s, W - 1) * SW + (hc ? c : jj + 1)) * 2;
            h[q] = get(p);
            l[q] = get(p + 1);
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int s = s0 + 8 * q;
            if (s < W && ncol > 0) ok = ok && tagged(h[q], tag) && tagged(l[q], tag);
          }
          if (__all(ok)) break;
          if (!keep_waiting(area, id, rounds)) {
            failed = true;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        if (failed) break;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (s0 + 8 * q < W) acc += join_d(h[q], l[q]);
      }
      if (failed) sfail = 1;
      if (hc) part[g][ci] = acc;
      __syncthreads();
      if (sfail) break;
      if (tid < ncol) {
        double t = part[0][tid];
#pragma unroll
        for (int q = 1; q < 8; ++q) t += part[q][tid];
        wvec[jj + 1 + tid] = t * (1.0 / tau);
      }
      __syncthreads();
      // a_jc -= w_c (row jj), a_rc += (-u_r) w_c (rows below)
#pragma unroll
      for (int c2 = 0; c2 < SW; ++c2) {
        if (c2 > jj && c2 < sw) {
          const double wc = wvec[c2];
          if (row == jj)
            v[c2] -= wc;
          else if (own && row > jj)
            v[c2] = fma(-u, wc, v[c2]);
        }
      }
    }
    if (w == 0 && tid == 0) taus[jj] = tau;
  }
  if (own) {
#pragma unroll
    for (int c = 0; c < SW; ++c)
      if (c < sw) A[(int64_t)row * lda + c] = v[c];
  }
}

// U (mrows x sw, ldu) = the unit lower part of the factored sub-panel P: 1 on the diagonal,
// u2 below, 0 above
__global__ __launch_bounds__(256) void form_u_kernel(int64_t mrows, int sw, const double* P,
                                                     int64_t ldp, double* U, int64_t ldu) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= mrows * sw) return;
  const int64_t r = e / sw;
  const int c = static_cast<int>(e % sw);
  U[r * ldu + c] = r > c ? P[r * ldp + c] : (r == c ? 1.0 : 0.0);
}

// T[i, i] = taus[i] for i < b (the diagonal of a block factor after U^T U)
__global__ __launch_bounds__(256) void set_tau_diag_kernel(int b, const double* taus, double* T,
                                                           int64_t ldt) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < b) T[(int64_t)i * ldt + i] = taus[i];
}

int launch_qr_panel(int64_t mrows, int sw, double* A, int64_t lda, double* taus, hipStream_t s) {
  const int64_t W = (mrows + ROWS - 1) / ROWS;
  if (W > WMAX) {
    set_error("QR panel of %lld rows exceeds the %d x %d rows one launch holds",
              (long long)mrows, WMAX, ROWS);
    return DDLA_ERR_UNSUPPORTED;
  }
  unsigned first = 0;
  DDLA_RETURN_IF(next_tag(WS_QR_SYNC, QR_AREA_BYTES, static_cast<unsigned>(sw) + 2u, s, &first));
  void* ws = nullptr;
  DDLA_RETURN_IF(workspace(WS_QR_SYNC, QR_AREA_BYTES, &ws));
  char* area = static_cast<char*>(ws);
  if (ctx().wait_limit_set)
    DDLA_RETURN_IF(launch_fill_u32(area + QR_ABORT_OFF + sizeof(unsigned), 1, ctx().wait_limit, s));
  ProfScope prof("qr_panel_f64", 4.0 * (double)mrows * sw * sw / 2.0,
                 16.0 * (double)mrows * sw, s);
  hipLaunchKernelGGL(qr_panel_kernel, dim3((unsigned)W), dim3(ROWS), 0, s, (int)mrows, sw, A, lda,
                     taus, reinterpret_cast<u64*>(area), first - 1u);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

unsigned* qr_abort_word() {
  void* area = ctx().ws[WS_QR_SYNC];
  return area ? reinterpret_cast<unsigned*>(static_cast<char*>(area) + QR_ABORT_OFF) : nullptr;
}

// C (mrows x nc, ldc) <- (I - U Tb^-1 U^T)^T C (trans) or (I - U Tb^-1 U^T) C; U mrows x b
// (ldu), Tb b x b upper (ldt), W >= b x nc scratch
int apply_block(bool trans, int64_t mrows, int b, const double* U, int64_t ldu, const double* Tb,
                int64_t ldt, double* C, int64_t ldc, int64_t nc, double* W, hipStream_t s) {
  if (nc <= 0 || mrows <= 0 || b <= 0) return 0;
  // W = U^T C
  DDLA_RETURN_IF(launch_dgemm(b, nc, mrows, 1.0, U, ldu, false, C, ldc, true, 0.0, W, nc, 0, s));
  // W = Tb^-T W or Tb^-1 W
  DDLA_RETURN_IF(trsm_left<double>(false, false, trans, b, nc, Tb, ldt, W, nc, s));
  // C -= U W
  return launch_dgemm(mrows, nc, b, -1.0, U, ldu, true, W, nc, true, 1.0, C, ldc, 0, s);
}

// the b x b factor of a block: strictly upper part of U^T U, tau on the diagonal
int block_t(int64_t mrows, int b, const double* U, int64_t ldu, const double* taus, double* Tb,
            int64_t ldt, hipStream_t s) {
  DDLA_RETURN_IF(launch_dgemm(b, b, mrows, 1.0, U, ldu, false, U, ldu, true, 0.0, Tb, ldt, 2, s));
  hipLaunchKernelGGL(set_tau_diag_kernel, dim3(ceil_div(b, 256)), dim3(256), 0, s, b, taus, Tb,
                     ldt);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

struct QrScratch {
  double* U;     // m x 256 explicit Householder vectors of the current block
  double* W;     // 256 x n products
  double* taus;  // min(m, n)
};

int qr_scratch(int64_t m, int64_t n, QrScratch* q) {
  const int64_t k = std::min(m, n);
  const size_t u = (size_t)m * QR_NB, w = (size_t)QR_NB * std::max<int64_t>(n, 1);
  void* ws = nullptr;
  DDLA_RETURN_IF(workspace(WS_QR_WORK, (u + w + (size_t)k + 64) * sizeof(double), &ws));
  q->U = static_cast<double*>(ws);
  q->W = q->U + u;
  q->taus = q->W + w;
  return 0;
}

}  // namespace

int qr_t_rows(int64_t m, int64_t n) {
  return static_cast<int>(std::min<int64_t>(QR_NB, std::max(m, n)));
}

// QRDecompose (QR.cs:130-149, :326-395)
int dgeqrf_internal(int64_t m, int64_t n, double* a, int64_t lda, double* t, int64_t ldt) {
  hipStream_t s = stream();
  const int64_t kmax = std::min(m, n);
  if (kmax == 0) return 0;
  if (m > (int64_t)WMAX * ROWS) {
    set_error("QR of %lld rows: at most %d rows are supported", (long long)m, WMAX * ROWS);
    return DDLA_ERR_UNSUPPORTED;
  }
  const int tb = qr_t_rows(m, n);
  QrScratch q;
  DDLA_RETURN_IF(qr_scratch(m, n, &q));
  for (int64_t k = 0; k < kmax; k += tb) {
    const int b = (int)std::min<int64_t>(tb, kmax - k);
    const int64_t mp = m - k;
    double* U = q.U;  // mp x b, ldu = b
    DDLA_RETURN_IF(launch_fill_u32(U, (size_t)mp * b * 2, 0u, s));
    for (int sp = 0; sp < b; sp += SW) {
      const int sw = std::min(SW, b - sp);
      const int64_t mpp = mp - sp;
      double* P = a + (k + sp) * lda + (k + sp);
      DDLA_RETURN_IF(launch_qr_panel(mpp, sw, P, lda, q.taus + k + sp, s));
      hipLaunchKernelGGL(form_u_kernel, dim3(ceil_div(mpp * sw, 256)), dim3(256), 0, s, mpp, sw,
                         P, lda, U + (int64_t)sp * b + sp, (int64_t)b);
      DDLA_HIP_CHECK(hipGetLastError());
      const int64_t rest = b - sp - sw;
      if (rest > 0) {
        double* Ts = t + (int64_t)sp * ldt + (k + sp);
        const double* Us = U + (int64_t)sp * b + sp;
        DDLA_RETURN_IF(block_t(mpp, sw, Us, b, q.taus + k + sp, Ts, ldt, s));
        DDLA_RETURN_IF(apply_block(true, mpp, sw, Us, b, Ts, ldt, P + sw, lda, rest, q.W, s));
      }
    }
    double* Tb = t + k;
    DDLA_RETURN_IF(block_t(mp, b, U, b, q.taus + k, Tb, ldt, s));
    const int64_t nc = n - k - b;
    if (nc > 0)
      DDLA_RETURN_IF(apply_block(true, mp, b, U, b, Tb, ldt, a + k * lda + k + b, lda, nc, q.W, s));
  }
  return check_abort(qr_abort_word(), s, "QR panel");
}

// ApplyQlhfc (trans: B <- Q^T B, QR.cs:423-464) / ApplyQlnfc (B <- Q B, :466-521), with the
// blocks of the factorization: forward for Q^T, backward for Q
int dqr_apply_internal(bool trans, int64_t m, int64_t n, const double* a, int64_t lda,
                       const double* t, int64_t ldt, int64_t nrhs, double* b, int64_t ldb) {
  hipStream_t s = stream();
  const int64_t kmax = std::min(m, n);
  if (kmax == 0 || nrhs == 0) return 0;
  const int tb = qr_t_rows(m, n);
  QrScratch q;
  DDLA_RETURN_IF(qr_scratch(m, nrhs, &q));
  const int64_t nblk = (kmax + tb - 1) / tb;
  for (int64_t i = 0; i < nblk; ++i) {
    const int64_t blk = trans ? i : nblk - 1 - i;
    const int64_t k = blk * tb;
    const int bw = (int)std::min<int64_t>(tb, kmax - k);
    const int64_t mp = m - k;
    hipLaunchKernelGGL(form_u_kernel, dim3(ceil_div(mp * bw, 256)), dim3(256), 0, s, mp, bw,
                       a + k * lda + k, lda, q.U, (int64_t)bw);
    DDLA_HIP_CHECK(hipGetLastError());
    DDLA_RETURN_IF(apply_block(trans, mp, bw, q.U, bw, t + k, ldt, b + k * ldb, ldb, nrhs, q.W, s));
  }
  return 0;
}

// SolveUncheck (QR.cs:397-409): Y = Q^T B (B kept), X = R^-1 Y[0:n]; m >= n
int dqr_solve_internal(int64_t m, int64_t n, int64_t nrhs, const double* a, int64_t lda,
                       const double* t, int64_t ldt, const double* b, int64_t ldb, double* x,
                       int64_t ldx) {
  hipStream_t s = stream();
  void* ws = nullptr;
  DDLA_RETURN_IF(workspace(WS_QR_Y, (size_t)m * nrhs * sizeof(double), &ws));
  double* y = static_cast<double*>(ws);
  DDLA_RETURN_IF(launch_dcopy2d(m, nrhs, b, ldb, 1, y, nrhs, s));
  DDLA_RETURN_IF(dqr_apply_internal(true, m, n, a, lda, t, ldt, nrhs, y, nrhs));
  DDLA_RETURN_IF(trsm_left<double>(false, false, false, n, nrhs, a, lda, y, nrhs, s));
  return launch_dcopy2d(n, nrhs, y, nrhs, 1, x, ldx, s);
}

}  // namespace ddla
