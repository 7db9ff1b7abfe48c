// Triangular solves with many right-hand sides, row-major device matrices.
//
// Replaces bli_dtrsm (DDLA/BLAS/Native/3m.cs:151-162) as called from PLUDecBlock
// (DDLA/Factorizations/LU.cs:332), PLUSolveInternel (LU.cs:442-463), CholeskyLowerBlock
// (DDLA/Factorizations/Cholesky.cs:156) and CholeskySolveUncheck (Cholesky.cs:210-226).
//
// Structure: the triangle is cut into 128-row diagonal blocks.  A diagonal block is
// solved by trsm_slab_kernel: one workgroup per 64 right-hand-side columns keeps the
// packed triangle and its 128 x 64 slab of B in LDS and substitutes row by row (lanes
// run along the slab's columns, so every LDS access is conflict-free or a broadcast).
// The off-diagonal coupling is a GEMM update on the matrix cores (launch_dgemm).
#include "ddla_internal.hpp"

#include <algorithm>

namespace ddla {
namespace {

constexpr int TB = 128;  // diagonal block size

// Solves op(T) X = B in place for one bs x bs triangle T (strides rs, cs) and the
// bs x ncols block B.  LOWER selects forward substitution.  Non-unit diagonals divide.
template <typename T, bool LOWER, bool UNIT>
__global__ __launch_bounds__(256) void trsm_slab_kernel(int bs, int64_t ncols,
                                                        const T* __restrict__ A, int64_t rs,
                                                        int64_t cs, T* __restrict__ B,
                                                        int64_t ldb) {
  __shared__ T Ls[TB * (TB + 1) / 2];
  __shared__ T S[TB * 64];
  const int tid = threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.x * 64;
  const int cw = (int)min<int64_t>(64, ncols - c0);
  // packed triangle: lower (r >= c) at r(r+1)/2 + c; upper (r <= c) stored as its
  // transpose, at c(c+1)/2 + r
  // one flat loop over the packed index so every lane's loads are independent
  const int npk = bs * (bs + 1) / 2;
  for (int p = tid; p < npk; p += 256) {
    int r = (int)((sqrtf(8.0f * p + 1.0f) - 1.0f) * 0.5f);
    while (r * (r + 1) / 2 > p) --r;
    while ((r + 1) * (r + 2) / 2 <= p) ++r;
    const int c = p - r * (r + 1) / 2;
    Ls[p] = LOWER ? A[r * rs + c * cs] : A[c * rs + r * cs];
  }
  for (int idx = tid; idx < bs * 64; idx += 256) {
    const int r = idx >> 6, cc = idx & 63;
    S[idx] = (cc < cw) ? B[(int64_t)r * ldb + c0 + cc] : static_cast<T>(0);
  }
  __syncthreads();
  const int cc = tid & 63, g = tid >> 6;
  // entry (r, i) of the triangle in packed form
  auto tri = [&](int r, int i) -> T {
    return LOWER ? Ls[r * (r + 1) / 2 + i] : Ls[i * (i + 1) / 2 + r];
  };
  // row r of the slab is owned by group g = r % 4; each step loads a batch of owned rows
  // before storing them (one LDS round trip per batch instead of one per row)
  constexpr int CH = 8;
  if (LOWER) {
    if (!UNIT && g == 0) S[cc] = S[cc] / tri(0, 0);
    __syncthreads();
    for (int i = 0; i < bs - 1; ++i) {
      const T x = S[i * 64 + cc];
      for (int k0 = 0; k0 < TB / 4; k0 += CH) {
        T v[CH], l[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k) {
          const int r = g + 4 * (k0 + k);
          v[k] = S[r * 64 + cc];
          l[k] = (r > i && r < bs) ? tri(r, i) : T(0);
        }
#pragma unroll
        for (int k = 0; k < CH; ++k) {
          const int r = g + 4 * (k0 + k);
          if (r > i && r < bs) {
            T w = fma(-l[k], x, v[k]);
            if (!UNIT && r == i + 1) w = w / tri(r, r);
            S[r * 64 + cc] = w;
          }
        }
      }
      __syncthreads();
    }
  } else {
    const int last = bs - 1;
    if (!UNIT && g == 0) S[last * 64 + cc] = S[last * 64 + cc] / tri(last, last);
    __syncthreads();
    for (int i = last; i > 0; --i) {
      const T x = S[i * 64 + cc];
      for (int k0 = 0; k0 < TB / 4; k0 += CH) {
        T v[CH], l[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k) {
          const int r = g + 4 * (k0 + k);
          v[k] = S[r * 64 + cc];
          // upper: row r < i couples through element (r, i) = stored transpose
          l[k] = (r < i) ? tri(r, i) : T(0);
        }
#pragma unroll
        for (int k = 0; k < CH; ++k) {
          const int r = g + 4 * (k0 + k);
          if (r < i) {
            T w = fma(-l[k], x, v[k]);
            if (!UNIT && r == i - 1) w = w / tri(r, r);
            S[r * 64 + cc] = w;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int idx = tid; idx < bs * 64; idx += 256) {
    const int r = idx >> 6, c = idx & 63;
    if (c < cw) B[(int64_t)r * ldb + c0 + c] = S[idx];
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
  dim3 grid(ceil_div(ncols, 64)), block(256);
  ProfScope prof(sizeof(T) == 8 ? "trsm_slab_f64" : "trsm_slab_f32", (double)bs * bs * ncols,
                 sizeof(T) * (2.0 * bs * ncols + 0.5 * bs * bs), s);
  if (lower && unit)
    hipLaunchKernelGGL((trsm_slab_kernel<T, true, true>), grid, block, 0, s, bs, ncols, A, rs, cs, B, ldb);
  else if (lower)
    hipLaunchKernelGGL((trsm_slab_kernel<T, true, false>), grid, block, 0, s, bs, ncols, A, rs, cs, B, ldb);
  else if (unit)
    hipLaunchKernelGGL((trsm_slab_kernel<T, false, true>), grid, block, 0, s, bs, ncols, A, rs, cs, B, ldb);
  else
    hipLaunchKernelGGL((trsm_slab_kernel<T, false, false>), grid, block, 0, s, bs, ncols, A, rs, cs, B, ldb);
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
template int launch_trsm_slab<double>(bool, bool, int, int64_t, const double*, int64_t, int64_t,
                                      double*, int64_t, hipStream_t);
template int launch_trsm_slab<float>(bool, bool, int, int64_t, const float*, int64_t, int64_t,
                                     float*, int64_t, hipStream_t);

}  // namespace ddla
