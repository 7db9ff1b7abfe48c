// Triangular solves with many right-hand sides, row-major device matrices.
//
// Replaces bli_dtrsm (DDLA/BLAS/Native/3m.cs:151-162) as called from PLUDecBlock
// (DDLA/Factorizations/LU.cs:332), PLUSolveInternel (LU.cs:442-463), CholeskyLowerBlock
// (DDLA/Factorizations/Cholesky.cs:156) and CholeskySolveUncheck (Cholesky.cs:210-226).
//
// Structure: the triangle is cut into 128-row diagonal blocks.  A diagonal block is
// solved by trsm_slab_kernel: the triangle sits in LDS, a wave keeps a few right-hand-side
// columns in registers (two positions per lane) and substitutes down them together.
// The off-diagonal coupling is a GEMM update on the matrix cores (launch_dgemm).
#include "ddla_internal.hpp"
#include "wave_ops.hpp"

#include <algorithm>

namespace ddla {
namespace {

constexpr int TB = 128;  // diagonal block size

// Solves op(T) X = B in place for one bs x bs triangle T (element (r, c) at A[r rs + c cs])
// and the bs x ncols block B.  LOWER selects forward substitution; non-unit diagonals divide.
//
// Everything runs in solve-order positions q (row q for LOWER, row bs-1-q otherwise).  A wave
// solves CW columns of B: lane l keeps positions l and l + 64 of each of its columns in
// registers.  Step k (ascending) finishes x_k in its owner lane k mod 64 (division by t_kk
// unless UNIT), broadcasts it with v_readlane and every lane with an unsolved position q > k
// subtracts t_qk x_k with one fma: per column the FMA chain in k order of a row-by-row
// solve.  The strictly lower couplings sit in LDS as Mt[q][k] (stride TB + 1: the column
// reads of one step hit distinct banks), zero where q <= k.
//
// PIV (unit lower only): B first takes a panel's row interchanges, given as the (dst <- src)
// list of laswp_list_kernel with rows relative to B's first row: the rows of the block are
// read from their sources, and the rows that leave the block (dst >= bs; their sources all
// lie inside the block) are written to their destinations by the wave that owns the columns,
// before the block itself is overwritten.
template <typename T, bool LOWER, bool UNIT, int CW, bool PIV = false>
__global__ __launch_bounds__(256) void trsm_slab_kernel(int bs, int64_t ncols,
                                                        const T* __restrict__ A, int64_t rs,
                                                        int64_t cs, T* __restrict__ B,
                                                        int64_t ldb,
                                                        const int* __restrict__ list = nullptr,
                                                        const int* __restrict__ count = nullptr) {
  __shared__ T Mt[TB][TB + 1];
  __shared__ T dg[TB];
  __shared__ int srcq[PIV ? TB : 1];
  __shared__ int lst[PIV ? 4 * TB : 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  auto row_of = [&](int q) { return LOWER ? q : bs - 1 - q; };
  const int64_t c0 = ((int64_t)blockIdx.x * 4 + wave) * CW;
  const bool kfast = cs == 1;  // consecutive threads read consecutive addresses
  constexpr int PER = TB * TB / 256, BATCH = sizeof(T) == 4 ? 32 : 16;
  int cnt = 0;
  T tv[PIV ? PER : 1];  // PIV: this thread's part of the triangle, in flight during the list work
  if constexpr (PIV) {
    // The count, the whole list area (4 TB words; the entries past 2 cnt are not used) and the
    // triangle are requested together: one memory round trip before the list is used, not
    // three in a row.
    static_assert(4 * TB == 2 * 256, "list area = two words per thread");
    const int c = *count;
    const int l0 = list[tid], l1 = list[tid + 256];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int p = i * 256 + tid;
      const int q = kfast ? p / TB : p % TB, k = kfast ? p % TB : p / TB;
      const bool in = q < bs && k <= q;
      tv[i] = A[in ? (int64_t)row_of(q) * rs + (int64_t)row_of(k) * cs : 0];
    }
    cnt = min(c, 2 * TB);
    lst[tid] = l0;
    lst[tid + 256] = l1;
    for (int t = tid; t < TB; t += 256) srcq[t] = t;
    __syncthreads();
    for (int e = tid; e < cnt; e += 256)
      if (lst[2 * e] < bs) srcq[lst[2 * e]] = lst[2 * e + 1];
    __syncthreads();
  }
  T x0[CW], x1[CW];  // positions lane and lane + 64 (read first: in flight during the staging)
  const int r0 = lane < bs ? row_of(lane) : 0, r1 = lane + 64 < bs ? row_of(lane + 64) : 0;
  const int s0 = PIV ? srcq[r0] : r0, s1 = PIV ? srcq[r1] : r1;
#pragma unroll
  for (int c = 0; c < CW; ++c) {
    const bool in = c0 + c < ncols;
    const bool i0 = in && lane < bs, i1 = in && lane + 64 < bs;
    x0[c] = B[i0 ? (int64_t)s0 * ldb + c0 + c : 0];
    x1[c] = B[i1 ? (int64_t)s1 * ldb + c0 + c : 0];
  }
  if constexpr (PIV) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int p = i * 256 + tid;
      const int q = kfast ? p / TB : p % TB, k = kfast ? p % TB : p / TB;
      const T x = (q < bs && k <= q) ? tv[i] : T(0);
      if (k == q) dg[q] = x;
      Mt[q][k] = k == q ? T(0) : x;
    }
  }
  // Otherwise staging in batches of 32 (FP64: 16) reads per thread, issued back to back
  // (branch-free: an out-of-range element reads A[0] and is replaced by zero), then written to
  // LDS.
#pragma unroll
  for (int b0 = 0; b0 < (PIV ? 0 : PER); b0 += BATCH) {
    T v[BATCH];
#pragma unroll
    for (int i = 0; i < BATCH; ++i) {
      const int p = (b0 + i) * 256 + tid;
      const int q = kfast ? p / TB : p % TB, k = kfast ? p % TB : p / TB;
      const bool in = q < bs && k <= q;
      v[i] = A[in ? (int64_t)row_of(q) * rs + (int64_t)row_of(k) * cs : 0];
    }
#pragma unroll
    for (int i = 0; i < BATCH; ++i) {
      const int p = (b0 + i) * 256 + tid;
      const int q = kfast ? p / TB : p % TB, k = kfast ? p % TB : p / TB;
      const T x = (q < bs && k <= q) ? v[i] : T(0);
      if (k == q) dg[q] = x;
      Mt[q][k] = k == q ? T(0) : x;
    }
  }
#pragma unroll
  for (int c = 0; c < CW; ++c) {
    const bool in = c0 + c < ncols;
    x0[c] = (in && lane < bs) ? x0[c] : T(0);
    x1[c] = (in && lane + 64 < bs) ? x1[c] : T(0);
  }
  __syncthreads();
  if (c0 >= ncols) return;  // after the barrier: every wave helped stage the triangle
  if constexpr (PIV) {
    // every read of the block's sources is complete (their values were consumed above)
    for (int e = lane; e < cnt; e += 64) {
      const int d = lst[2 * e], sr = lst[2 * e + 1];
      if (d >= bs) {
#pragma unroll
        for (int c = 0; c < CW; ++c)
          if (c0 + c < ncols) B[(int64_t)d * ldb + c0 + c] = B[(int64_t)sr * ldb + c0 + c];
      }
    }
  }
  const int h = min(bs, 64);
#pragma unroll 2
  for (int k = 0; k < h; ++k) {
    const T m0 = Mt[lane][k], m1 = Mt[lane + 64][k];
    if (!UNIT && lane == k) {
      const T d = dg[k];
#pragma unroll
      for (int c = 0; c < CW; ++c) x0[c] = x0[c] / d;
    }
#pragma unroll
    for (int c = 0; c < CW; ++c) {
      const T xk = read_lane(x0[c], k);
      const T u = fma(-m0, xk, x0[c]);
      x0[c] = lane > k ? u : x0[c];
      x1[c] = fma(-m1, xk, x1[c]);
    }
  }
#pragma unroll 2
  for (int k = 64; k < bs; ++k) {
    const T m1 = Mt[lane + 64][k];
    if (!UNIT && lane == k - 64) {
      const T d = dg[k];
#pragma unroll
      for (int c = 0; c < CW; ++c) x1[c] = x1[c] / d;
    }
#pragma unroll
    for (int c = 0; c < CW; ++c) {
      const T xk = read_lane(x1[c], k - 64);
      const T u = fma(-m1, xk, x1[c]);
      x1[c] = lane + 64 > k ? u : x1[c];
    }
  }
#pragma unroll
  for (int c = 0; c < CW; ++c) {
    if (c0 + c >= ncols) break;
    if (lane < bs) B[(int64_t)r0 * ldb + c0 + c] = x0[c];
    if (lane + 64 < bs) B[(int64_t)r1 * ldb + c0 + c] = x1[c];
  }
}

}  // namespace

template <typename T>
int launch_trsm_slab(bool lower, bool unit, int bs, int64_t ncols, const T* A, int64_t rs,
                     int64_t cs, T* B, int64_t ldb, hipStream_t s) {
  if (ncols <= 0 || bs <= 0) return 0;
  if (bs > TB) {
    set_error("trsm block %d > %d", bs, TB);
    return DDLA_ERR_INVALID_ARG;
  }
  ProfScope prof(sizeof(T) == 8 ? "trsm_slab_f64" : "trsm_slab_f32", (double)bs * bs * ncols,
                 sizeof(T) * (2.0 * bs * ncols + 0.5 * bs * bs), s);
  // 4 columns per wave; 8 for wide blocks (half the triangle stagings per column)
  const bool wide = ncols >= 4096;
  const dim3 grid((unsigned)ceil_div(ncols, 4 * (wide ? 8 : 4))), block(256);
#define TRSM_SLAB(L, U)                                                                          \
  do {                                                                                           \
    if (wide)                                                                                    \
      hipLaunchKernelGGL((trsm_slab_kernel<T, L, U, 8>), grid, block, 0, s, bs, ncols, A, rs, cs, \
                         B, ldb);                                                                \
    else                                                                                         \
      hipLaunchKernelGGL((trsm_slab_kernel<T, L, U, 4>), grid, block, 0, s, bs, ncols, A, rs, cs, \
                         B, ldb);                                                                \
  } while (0)
  if (lower && unit) TRSM_SLAB(true, true);
  else if (lower) TRSM_SLAB(true, false);
  else if (unit) TRSM_SLAB(false, true);
  else TRSM_SLAB(false, false);
#undef TRSM_SLAB
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

// op(A) X = B, A n x n row-major (lda), B n x nrhs row-major (ldb); lower/upper refer to
// the stored triangle of A, transA solves with its transpose.
namespace {
inline int gemm_t(int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                  bool ak, const double* B, int64_t ldb, bool bn, double beta, double* C,
                  int64_t ldc, hipStream_t s) {
  return launch_dgemm(m, n, k, alpha, A, lda, ak, B, ldb, bn, beta, C, ldc, 0, s);
}
inline int gemm_t(int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                  bool ak, const float* B, int64_t ldb, bool bn, float beta, float* C,
                  int64_t ldc, hipStream_t s) {
  return launch_sgemm(m, n, k, alpha, A, lda, ak, B, ldb, bn, beta, C, ldc, 0, s);
}
}  // namespace

template <typename T>
int trsm_left(bool lower, bool unit, bool transA, int64_t n, int64_t nrhs, const T* A,
              int64_t lda, T* B, int64_t ldb, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return 0;
  // element (i,k) of op(A): (rs, cs) strides
  const int64_t rs = transA ? 1 : lda, cs = transA ? lda : 1;
  const bool low = transA ? !lower : lower;  // triangle of op(A)
  const bool a_kcontig = !transA;            // op(A)(i,k) at A[i*lda+k] when not transposed
  if (low) {
    for (int64_t k0 = 0; k0 < n; k0 += TB) {
      const int bs = (int)std::min<int64_t>(TB, n - k0);
      const T* Akk = A + k0 * rs + k0 * cs;
      DDLA_RETURN_IF(launch_trsm_slab<T>(true, unit, bs, nrhs, Akk, rs, cs, B + k0 * ldb, ldb, s));
      const int64_t rest = n - k0 - bs;
      if (rest > 0) {
        const T* Ark = A + (k0 + bs) * rs + k0 * cs;  // op(A)[k0+bs:, k0:k0+bs]
        DDLA_RETURN_IF(gemm_t(rest, nrhs, bs, T(-1), Ark, lda, a_kcontig, B + k0 * ldb, ldb, true,
                              T(1), B + (k0 + bs) * ldb, ldb, s));
      }
    }
  } else {
    int64_t end = n;
    while (end > 0) {
      const int bs = (int)std::min<int64_t>(TB, end);
      const int64_t k0 = end - bs;
      const T* Akk = A + k0 * rs + k0 * cs;
      DDLA_RETURN_IF(launch_trsm_slab<T>(false, unit, bs, nrhs, Akk, rs, cs, B + k0 * ldb, ldb, s));
      if (k0 > 0) {
        const T* Ark = A + 0 * rs + k0 * cs;  // op(A)[0:k0, k0:k0+bs]
        DDLA_RETURN_IF(gemm_t(k0, nrhs, bs, T(-1), Ark, lda, a_kcontig, B + k0 * ldb, ldb, true,
                              T(1), B, ldb, s));
      }
      end = k0;
    }
  }
  return 0;
}

template int trsm_left<double>(bool, bool, bool, int64_t, int64_t, const double*, int64_t,
                               double*, int64_t, hipStream_t);
template int trsm_left<float>(bool, bool, bool, int64_t, int64_t, const float*, int64_t, float*,
                              int64_t, hipStream_t);
// The row interchanges of a panel (list, *count pairs) followed by the unit-lower solve
// L11^-1 on the bs x ncols block B (rows relative to the panel's first row); replaces
// laswp_list + trsm_slab on the same columns with one launch.
template <typename T>
int launch_laswp_trsm_slab(int bs, int64_t ncols, const T* L, int64_t ldl, T* B, int64_t ldb,
                           const int* list, const int* count, hipStream_t s) {
  if (ncols <= 0 || bs <= 0) return 0;
  if (bs > TB) {
    set_error("trsm block %d > %d", bs, TB);
    return DDLA_ERR_INVALID_ARG;
  }
  ProfScope prof(sizeof(T) == 8 ? "laswp_trsm_slab_f64" : "laswp_trsm_slab_f32",
                 (double)bs * bs * ncols, sizeof(T) * (2.0 * bs * ncols + 0.5 * bs * bs), s);
  // few columns (the mixed LU's next block: 128): one column per wave, four times the
  // workgroups (measured 22 us against 28 us with four columns per wave at n = 16384)
  const int cw = ncols >= 4096 ? 8 : ncols <= 512 ? 1 : 4;
  const dim3 grid((unsigned)ceil_div(ncols, 4 * cw)), block(256);
  if (cw == 8)
    hipLaunchKernelGGL((trsm_slab_kernel<T, true, true, 8, true>), grid, block, 0, s, bs, ncols,
                       L, ldl, (int64_t)1, B, ldb, list, count);
  else if (cw == 1)
    hipLaunchKernelGGL((trsm_slab_kernel<T, true, true, 1, true>), grid, block, 0, s, bs, ncols,
                       L, ldl, (int64_t)1, B, ldb, list, count);
  else
    hipLaunchKernelGGL((trsm_slab_kernel<T, true, true, 4, true>), grid, block, 0, s, bs, ncols,
                       L, ldl, (int64_t)1, B, ldb, list, count);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

template int launch_laswp_trsm_slab<double>(int, int64_t, const double*, int64_t, double*,
                                            int64_t, const int*, const int*, hipStream_t);
template int launch_laswp_trsm_slab<float>(int, int64_t, const float*, int64_t, float*, int64_t,
                                           const int*, const int*, hipStream_t);
template int launch_trsm_slab<double>(bool, bool, int, int64_t, const double*, int64_t, int64_t,
                                      double*, int64_t, hipStream_t);
template int launch_trsm_slab<float>(bool, bool, int, int64_t, const float*, int64_t, int64_t,
                                     float*, int64_t, hipStream_t);

}  // namespace ddla
