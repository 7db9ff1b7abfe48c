// Triangular solves with many right-hand sides, row-major device matrices.
//
// Replaces bli_dtrsm (DDLA/BLAS/Native/3m.cs:151-162) as called from PLUDecBlock
// (DDLA/Factorizations/LU.cs:332), PLUSolveInternel (LU.cs:442-463), CholeskyLowerBlock
// (DDLA/Factorizations/Cholesky.cs:156) and CholeskySolveUncheck (Cholesky.cs:210-226).
//
// Structure: the triangle is cut into 128-row diagonal blocks.  A diagonal block is
// solved by trsm_slab_kernel: the packed triangle sits in LDS and each lane substitutes
// down one right-hand-side column on its own (no barriers).
// The off-diagonal coupling is a GEMM update on the matrix cores (launch_dgemm).
#include "ddla_internal.hpp"

#include <algorithm>

namespace ddla {
namespace {

constexpr int TB = 128;  // diagonal block size

// Solves op(T) X = B in place for one bs x bs triangle T (strides rs, cs) and the bs x ncols
// block B.  LOWER selects forward substitution; non-unit diagonals divide.
//
// Each lane owns one column of B and solves it alone, so there is no barrier after the
// triangle is staged in LDS.  Everything runs in solve-order positions q (row q for LOWER,
// row bs-1-q otherwise): M[qk][q] is the coupling of the already solved position qk into
// position q > qk, so the couplings a step needs are contiguous (16-byte LDS reads, the
// same address in every lane: broadcasts) and M[q][q] is the diagonal.  Positions go in
// sub-blocks of SB: the sub-block's values sit in registers, take the contributions of the
// positions already solved (read back from B, in order), then the substitution inside; each
// x_q = (b_q - sum_k t_qk x_k) / t_qq is the FMA chain in k order of a row-by-row solve.
template <typename T, bool LOWER, bool UNIT>
__global__ __launch_bounds__(256) void trsm_slab_kernel(int bs, int64_t ncols,
                                                        const T* __restrict__ A, int64_t rs,
                                                        int64_t cs, T* __restrict__ B,
                                                        int64_t ldb) {
  constexpr int SB = 16;
  __shared__ __attribute__((aligned(16))) T M[TB][TB];
  const int tid = threadIdx.x;
  auto row_of = [&](int q) { return LOWER ? q : bs - 1 - q; };
  for (int p = tid; p < bs * bs; p += 256) {
    const int qk = p / bs, q = p % bs;  // q runs along memory for the lower case
    const int r = row_of(q), c = row_of(qk);
    // element (r, c) of the stored triangle: lower r >= c, upper r <= c
    M[qk][q] = (q >= qk) ? A[(int64_t)r * rs + (int64_t)c * cs] : T(0);
  }
  __syncthreads();
  const int64_t col = (int64_t)blockIdx.x * 256 + tid;
  if (col >= ncols) return;
  T* b = B + col;
  constexpr int VW = 16 / sizeof(T);  // elements per 16-byte read
  typedef T vec __attribute__((ext_vector_type(VW)));
  for (int s0 = 0; s0 < bs; s0 += SB) {
    T x[SB];
#pragma unroll
    for (int j = 0; j < SB; ++j) x[j] = (s0 + j < bs) ? b[(int64_t)row_of(s0 + j) * ldb] : T(0);
    // positions already solved, in order
#pragma unroll 2
    for (int qk = 0; qk < s0; ++qk) {
      const T xk = b[(int64_t)row_of(qk) * ldb];
#pragma unroll
      for (int v = 0; v < SB / VW; ++v) {
        const vec m = *reinterpret_cast<const vec*>(&M[qk][s0 + v * VW]);
#pragma unroll
        for (int e = 0; e < VW; ++e) x[v * VW + e] = fma(-m[e], xk, x[v * VW + e]);
      }
    }
    // inside the sub-block (entries past bs are zero couplings of zero values)
#pragma unroll
    for (int i = 0; i < SB; ++i) {
      if (!UNIT && s0 + i < bs) x[i] = x[i] / M[s0 + i][s0 + i];
#pragma unroll
      for (int j = i + 1; j < SB; ++j) x[j] = fma(-M[s0 + i][s0 + j], x[i], x[j]);
    }
#pragma unroll
    for (int j = 0; j < SB; ++j)
      if (s0 + j < bs) b[(int64_t)row_of(s0 + j) * ldb] = x[j];
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
  dim3 grid(ceil_div(ncols, 256)), block(256);
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
