// Kernels of the blocked Cholesky (lower).  Reference: CholeskyLowerBlock /
// CholeskyLowerUnblock / Sqrt (DDLA/Factorizations/Cholesky.cs:127-160,228-236).
//
//   potf2_rec_kernel  the nb x nb diagonal block in LDS, in 16-column blocks: the diagonal
//                     16 x 16 in one wave's registers, substitution and rank-16 update by all
//                     (16 measured faster than 32 and 8: the single-wave diagonal step grows
//                     with the block width, the number of barriers with its inverse)
//   trsm_rlt_kernel   A21 <- A21 * L11^-T (or * U11^-1), four lanes per row of A21 with the
//                     row in registers, the triangle in LDS
//   transpose_kernel  in-place transpose of a square matrix (the upper factorization is the
//                     lower one of the transposed storage)
#include "ddla_internal.hpp"

#include <utility>

namespace ddla {
namespace {

constexpr int NB = 128;


// LU without pivoting of one nb x nb block in LDS (LUDecUnblock, LU.cs:188-203): for each
// column i, a21 *= 1/a11 (InvScal multiplies by the reciprocal) and A22 -= a21 a12^T.
// Each step's rank-1 update issues the LDS loads of a batch of rows before its stores.
// A zero pivot sets *info = info_base + i + 1 (first failure wins) and stops.
template <typename T>
__global__ __launch_bounds__(256) void getf2_nopiv_kernel(int nb, T* __restrict__ A, int64_t lda,
                                                          int32_t* __restrict__ info,
                                                          int info_base) {
  __shared__ T t[NB * (NB + 1)];
  constexpr int LD = NB + 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
  for (int b0 = 0; b0 < NB * NB / 256; b0 += 16) {  // 16 reads in flight per thread
    T v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int idx = (b0 + i) * 256 + tid;
      v[i] = A[idx < nb * nb ? (int64_t)(idx / nb) * lda + idx % nb : 0];
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int idx = (b0 + i) * 256 + tid;
      if (idx < nb * nb) t[(idx / nb) * LD + idx % nb] = v[i];
    }
  }
  __syncthreads();
  bool stop = false;
  for (int i = 0; i < nb; ++i) {
    const T pv = t[i * LD + i];
    if (pv == T(0)) {
      if (tid == 0) atomicCAS(info, 0, info_base + i + 1);
      stop = true;
      break;
    }
    const T rinv = T(1) / pv;
    const int c0 = i + 1 + lane, c1 = c0 + 64;
    const bool h0 = c0 < nb, h1 = c1 < nb;
    const T u0 = h0 ? t[i * LD + c0] : T(0), u1 = h1 ? t[i * LD + c1] : T(0);
    constexpr int CH = 8;
    for (int k0 = 0; k0 < NB / 4; k0 += CH) {
      T l[CH], a0[CH], a1[CH];
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const T* tr = t + (wave + 4 * (k0 + k)) * LD;
        l[k] = tr[i];
        a0[k] = h0 ? tr[c0] : T(0);
        a1[k] = h1 ? tr[c1] : T(0);
      }
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int r = wave + 4 * (k0 + k);
        if (r > i && r < nb) {
          T* tr = t + r * LD;
          const T lk = l[k] * rinv;
          if (h0) tr[c0] = fma(-lk, u0, a0[k]);
          if (h1) tr[c1] = fma(-lk, u1, a1[k]);
          if (lane == 0) tr[i] = lk;
        }
      }
    }
    __syncthreads();
  }
  (void)stop;
  for (int idx = tid; idx < nb * nb; idx += 256) {
    const int r = idx / nb, c = idx % nb;
    A[(int64_t)r * lda + c] = t[r * LD + c];
  }
}

// Lower Cholesky of one nb x nb diagonal block (nb <= 128) in LDS, blocked by 32 columns
// (CholeskyLowerUnblock, Cholesky.cs:127-140, reorganised; every element takes the same fmas in
// the same column order, and the same scaling by the reciprocal of the pivot, as the unblocked
// loop):
//   (a) the 32 x 32 diagonal block, in wave 0's registers (lane r = row r), pivot columns
//       broadcast with v_readlane: no barrier per column;
//   (b) the rows below it: row-wise substitution against the factored block (one thread per
//       row, the block in LDS);
//   (c) the trailing lower triangle takes the block's rank-32 update (one element per thread
//       at a time, fma over the block's columns in order).
// *info = info_base + first non-positive pivot + 1 (if *info was 0); the block is then left as
// it stands.
template <typename T>
__device__ __forceinline__ T rl(T x, int lane) {
  if constexpr (sizeof(T) == 8) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
    const unsigned hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __builtin_bit_cast(T, (static_cast<unsigned long long>(hi) << 32) | lo);
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), lane));
  }
}

template <typename T>
__global__ __launch_bounds__(256) void potf2_rec_kernel(int nb, T* __restrict__ A, int64_t lda,
                                                        int32_t* __restrict__ info, int info_base) {
  constexpr int CB = 16;
  __shared__ T S[NB][NB + 1];
  __shared__ T piv_s[NB], piv_r[NB];  // sqrt(a_kk) and its reciprocal
  __shared__ int stop;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
  // fixed trip count, unrolled: a thread's reads are in flight together
  for (int b0 = 0; b0 < NB * NB / 256; b0 += 16) {  // 16 reads in flight per thread
    T v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int e = (b0 + i) * 256 + tid;
      const int r = e / nb, c = e % nb;
      v[i] = A[(e < nb * nb && c <= r) ? (int64_t)r * lda + c : 0];
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int e = (b0 + i) * 256 + tid;
      const int r = e / nb, c = e % nb;
      if (e < nb * nb && c <= r) S[r][c] = v[i];
    }
  }
  if (tid == 0) stop = 0;
  __syncthreads();
  for (int j0 = 0; j0 < nb; j0 += CB) {
    const int bs = min(CB, nb - j0);
    // (a) diagonal block in wave 0
    if (wave == 0) {
      T v[CB];
#pragma unroll
      for (int c = 0; c < CB; ++c) v[c] = (lane < bs && c <= lane && c < bs) ? S[j0 + lane][j0 + c] : T(0);
      int st = 0;
#pragma unroll
      for (int k = 0; k < CB; ++k) {
        if (k < bs && st == 0) {  // uniform
          const T d = rl(v[k], k);
          if (!(d > T(0))) {
            st = j0 + k + 1;
          } else {
            const T sq = sqrt(d);
            const T rinv = T(1) / sq;
            if (lane == k) v[k] = sq;
            else if (lane > k && sq != T(1)) v[k] *= rinv;
            if (lane == 0) {
              piv_s[j0 + k] = sq;
              piv_r[j0 + k] = rinv;
            }
            const T xr = -v[k];
#pragma unroll
            for (int c = k + 1; c < CB; ++c) {
              const T lck = rl(v[k], c);  // row c's multiplier for column k
              if (lane >= c && lane < bs) v[c] = fma(lck, xr, v[c]);
            }
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      if (st == 0) {
#pragma unroll
        for (int c = 0; c < CB; ++c)
          if (lane < bs && c <= lane) S[j0 + lane][j0 + c] = v[c];
      } else if (lane == 0) {
        stop = st;
      }
    }
    __syncthreads();
    if (stop) break;  // uniform
    const int i0 = j0 + CB, m2 = nb - i0;  // rows below the block
    if (m2 <= 0) break;
    // (b) rows below: x_k = a_k / l_kk (reciprocal), then x_c -= l_ck x_k for c > k
    if (tid < m2) {
      const int i = i0 + tid;
      T x[CB];
#pragma unroll
      for (int c = 0; c < CB; ++c) x[c] = S[i][j0 + c];
#pragma unroll
      for (int k = 0; k < CB; ++k) {
        x[k] *= piv_r[j0 + k];  // exact when l_kk = 1 (reciprocal 1)
        const T xr = -x[k];
#pragma unroll
        for (int c = k + 1; c < CB; ++c) x[c] = fma(S[j0 + c][j0 + k], xr, x[c]);
        // keep each step's LDS reads next to their use (the compiler would otherwise issue
        // all 496 of them up front)
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int c = 0; c < CB; ++c) S[i][j0 + c] = x[c];
    }
    __syncthreads();
    // (c) trailing lower triangle: a_ic -= sum_p l_ip l_cp, p over the block in order; the
    // m2 (m2 + 1) / 2 elements with cc <= ii enumerated row by row (no idle lanes for the
    // upper half)
    for (int e = tid; e < m2 * (m2 + 1) / 2; e += 256) {
      int ii = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
      while ((ii + 1) * (ii + 2) / 2 <= e) ++ii;
      while (ii * (ii + 1) / 2 > e) --ii;
      const int cc = e - ii * (ii + 1) / 2;
      const int i = i0 + ii, c = i0 + cc;
      T x = S[i][c];
#pragma unroll
      for (int p = 0; p < CB; ++p) x = fma(S[c][j0 + p], -S[i][j0 + p], x);
      S[i][c] = x;
    }
    __syncthreads();
  }
  if (stop) {
    if (tid == 0) atomicCAS(info, 0, info_base + stop);
    return;
  }
  for (int e = tid; e < nb * nb; e += 256) {
    const int r = e / nb, c = e % nb;
    if (c <= r) A[(int64_t)r * lda + c] = S[r][c];
  }
}

// B (nrows x nb) <- B * U^-1 for an upper nb x nb U (non-unit), nb <= 128: TL = U stored as
// it is (U[k][c] at L[k ldl + c]); otherwise U = L^T for the lower L stored at L (B <- B L^-T).
// Four lanes per row of B (a DPP quad): lane q holds the column groups 4 (q + 4 t) .. + 3,
// t < 8, in registers.  Column k is finished by its owner (x = a_k times the reciprocal of
// u_kk, the inverted-diagonal form of BLIS' TRSM kernels), broadcast to the
// quad, and every lane subtracts x U[k][c] from its columns c > k with one fma each (U's
// strictly upper part in LDS, zero beyond nb; 16-byte broadcast reads).  Per row this is the
// right-looking substitution, k ascending.
template <int K, typename T>
__device__ __forceinline__ T quad_bcast(T x) {
  constexpr int o = (K >> 2) & 3;  // owner lane of column K within the quad
  constexpr int ctrl = o | (o << 2) | (o << 4) | (o << 6);
  if constexpr (sizeof(T) == 8) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b), ctrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), ctrl, 0xf, 0xf, false);
    return __builtin_bit_cast(T, (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                                     static_cast<unsigned>(lo));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), ctrl, 0xf,
                                                          0xf, false));
  }
}

template <typename T>
using Quad4 = T __attribute__((ext_vector_type(4)));

template <int K, typename T>
__device__ __forceinline__ void rlt_step(Quad4<T> (&a)[8], const Quad4<T> (*U)[NB / 4],
                                         const T* rinv, int q) {
  constexpr int TT = K / 16, E = K % 4;
  T x = T(0);
  if (q == ((K >> 2) & 3)) {
    x = a[TT][E] * rinv[K];
    a[TT][E] = x;
  }
  x = quad_bcast<K>(x);
#pragma unroll
  for (int t = TT; t < 8; ++t) {  // groups that can hold a column > K
    const Quad4<T> u = U[K][q + 4 * t];
#pragma unroll
    for (int e = 0; e < 4; ++e) a[t][e] = fma(-x, u[e], a[t][e]);
  }
  // keep each step's LDS reads next to their use (hoisting them across steps only raises
  // register pressure)
  __builtin_amdgcn_sched_barrier(0);
}

template <typename T, int... Ks>
__device__ __forceinline__ void rlt_steps(std::integer_sequence<int, Ks...>, Quad4<T> (&a)[8],
                                          const Quad4<T> (*U)[NB / 4], const T* rinv, int q) {
  (rlt_step<Ks>(a, U, rinv, q), ...);
}

template <typename T, bool TL>
__global__ __launch_bounds__(256) void trsm_rlt_kernel(int nb, int64_t nrows,
                                                       const T* __restrict__ L, int64_t ldl,
                                                       T* __restrict__ B, int64_t ldb,
                                                       bool unit) {
  __shared__ Quad4<T> U[NB][NB / 4];  // strictly upper part of U, zero beyond nb
  __shared__ T rinv[NB];  // 1 / u_kk (1 beyond nb, and everywhere for a unit diagonal)
  const int tid = threadIdx.x;
  const T dv = (tid < nb && !unit) ? L[(int64_t)tid * ldl + tid] : T(1);  // divided later
  // batches of 16 branch-free reads per thread in flight together
  for (int b0 = 0; b0 < NB * NB / 256; b0 += 16) {
    T v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int e = (b0 + i) * 256 + tid, r = e / NB, c = e % NB;
      const bool in = r < nb && c < nb && c > r;
      v[i] = L[in ? (TL ? (int64_t)r * ldl + c : (int64_t)c * ldl + r) : 0];
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int e = (b0 + i) * 256 + tid, r = e / NB, c = e % NB;
      reinterpret_cast<T*>(&U[0][0])[e] = (r < nb && c < nb && c > r) ? v[i] : T(0);
    }
  }
  if (tid < NB) rinv[tid] = T(1) / dv;
  __syncthreads();
  const int q = tid & 3;
  const int64_t row = (int64_t)blockIdx.x * 64 + (tid >> 2);
  const bool own = row < nrows;
  T* br = B + (own ? row : nrows - 1) * ldb;
  Quad4<T> a[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int c = 4 * (q + 4 * t);
#pragma unroll
    for (int e = 0; e < 4; ++e) a[t][e] = c + e < nb ? br[c + e] : T(0);
  }
  rlt_steps(std::make_integer_sequence<int, NB>{}, a, U, rinv, q);
  if (own) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int c = 4 * (q + 4 * t);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (c + e < nb) br[c + e] = a[t][e];
    }
  }
}

// in-place transpose of an n x n row-major matrix: workgroup (bi, bj), bi <= bj, swaps
// the 32 x 32 tiles (bi, bj) and (bj, bi) through LDS
__global__ __launch_bounds__(256) void transpose_kernel(int64_t n, double* __restrict__ a,
                                                        int64_t lda) {
  __shared__ double t1[32][33], t2[32][33];
  const int64_t bi = blockIdx.y, bj = blockIdx.x;
  if (bi > bj) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int64_t r1 = bi * 32 + k, c1 = bj * 32 + tx;
    t1[k][tx] = (r1 < n && c1 < n) ? a[r1 * lda + c1] : 0.0;
    const int64_t r2 = bj * 32 + k, c2 = bi * 32 + tx;
    t2[k][tx] = (r2 < n && c2 < n) ? a[r2 * lda + c2] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int64_t r1 = bi * 32 + k, c1 = bj * 32 + tx;
    if (r1 < n && c1 < n) a[r1 * lda + c1] = t2[tx][k];
    const int64_t r2 = bj * 32 + k, c2 = bi * 32 + tx;
    if (bi != bj && r2 < n && c2 < n) a[r2 * lda + c2] = t1[tx][k];
  }
}

// LDL^T of one nb x nb diagonal block, lower triangle (LDLTLowerUnblock, LDLT.cs:141-158): for
// each column i with pivot d = a_ii, l = a21 * (1/d) (InvScaled multiplies by the reciprocal)
// and the lower triangle of A22 gets a_rc += (-d l_r) l_c.  The block's lower triangle sits in
// LDS; column i of the block is also kept in a double-buffered LDS vector, written by the
// previous step, so one barrier per column suffices.  Thread (wave, lane) updates rows
// wave + 4k and columns i + 1 + lane, i + 65 + lane.  Only the lower triangle (with the
// diagonal) is read and written back.  A zero pivot sets *info = info_base + i + 1 (first
// failure wins) and leaves the block as it stands.
__global__ __launch_bounds__(256) void ldlt_diag_kernel(int nb, double* __restrict__ A,
                                                        int64_t lda, int32_t* __restrict__ info,
                                                        int info_base) {
  constexpr int LD = NB + 1;
  __shared__ double t[NB * LD];
  __shared__ double colv[2][NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
  for (int b0 = 0; b0 < NB * NB / 256; b0 += 16) {  // 16 reads in flight per thread
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int idx = (b0 + q) * 256 + tid, r = idx / NB, c = idx % NB;
      v[q] = A[(r < nb && c <= r) ? (int64_t)r * lda + c : 0];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int idx = (b0 + q) * 256 + tid, r = idx / NB, c = idx % NB;
      if (r < nb && c <= r) {
        t[r * LD + c] = v[q];
        if (c == 0) colv[0][r] = v[q];
      }
    }
  }
  __syncthreads();
  for (int i = 0; i < nb; ++i) {
    const int par = i & 1;
    const double d = colv[par][i];
    if (d == 0.0) {
      if (tid == 0) atomicCAS(info, 0, info_base + i + 1);
      break;
    }
    const double rinv = 1.0 / d;
    const int c0 = i + 1 + lane, c1 = c0 + 64;
    const bool h0 = c0 < nb, h1 = c1 < nb;
    const double l0 = h0 ? colv[par][c0] * rinv : 0.0;
    const double l1 = h1 ? colv[par][c1] * rinv : 0.0;
    constexpr int CH = 8;
    // rows wave + 4k > i start at k = (i - wave) / 4 + 1; chunks below hold no such row
    const int kmin = i >= wave ? (i - wave) / 4 + 1 : 0;
    for (int k0 = kmin / CH * CH; k0 < NB / 4; k0 += CH) {
      double a0[CH], a1[CH], x[CH];
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int r = wave + 4 * (k0 + k);
        const double* tr = t + r * LD;
        x[k] = colv[par][r];
        a0[k] = (h0 && c0 <= r) ? tr[c0] : 0.0;
        a1[k] = (h1 && c1 <= r) ? tr[c1] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int r = wave + 4 * (k0 + k);
        if (r > i && r < nb) {
          double* tr = t + r * LD;
          const double lr = x[k] * rinv;
          const double ar = -d * lr;  // alpha x_r with alpha = -d
          if (h0 && c0 <= r) {
            const double v = fma(ar, l0, a0[k]);
            tr[c0] = v;
            if (c0 == i + 1) colv[par ^ 1][r] = v;  // the next column, for the next step
          }
          if (h1 && c1 <= r) tr[c1] = fma(ar, l1, a1[k]);
          if (lane == 0) tr[i] = lr;
        }
      }
    }
    __syncthreads();
  }
  for (int idx = tid; idx < nb * nb; idx += 256) {
    const int r = idx / nb, c = idx % nb;
    if (c <= r) A[(int64_t)r * lda + c] = t[r * LD + c];
  }
}

// Y = A21 (rows x b) and A21[:, j] *= 1 / d_j with d_j = A11[j, j] (LDLT.cs:188-191)
__global__ __launch_bounds__(256) void ldlt_scale_kernel(int64_t rows, int b,
                                                         const double* __restrict__ A11,
                                                         int64_t lda, double* __restrict__ A21,
                                                         double* __restrict__ Y, int64_t ldy) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= rows * b) return;
  const int64_t r = e / b;
  const int c = static_cast<int>(e % b);
  const double v = A21[r * lda + c];
  Y[r * ldy + c] = v;
  A21[r * lda + c] = v * (1.0 / A11[(int64_t)c * lda + c]);
}

// X[i, :] *= 1 / A[i, i] (LDLT.cs:274-275)
__global__ __launch_bounds__(256) void scale_rows_by_inv_diag_kernel(int64_t n, int64_t nrhs,
                                                                     const double* __restrict__ A,
                                                                     int64_t lda,
                                                                     double* __restrict__ X,
                                                                     int64_t ldx) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n * nrhs) return;
  const int64_t r = e / nrhs, c = e % nrhs;
  X[r * ldx + c] *= 1.0 / A[r * lda + r];
}

}  // namespace

template <typename T>
int launch_potf2_ex(int nb, T* A, int64_t lda, int32_t* info, int info_base, hipStream_t s) {
  if (nb > NB) {
    set_error("potf2 block %d > %d", nb, NB);
    return DDLA_ERR_INVALID_ARG;
  }
  if (nb <= 0) return 0;
  ProfScope prof(sizeof(T) == 8 ? "potf2_f64" : "potf2_f32", (double)nb * nb * nb / 3.0,
                 sizeof(T) * (double)nb * nb, s);
  hipLaunchKernelGGL(potf2_rec_kernel<T>, dim3(1), dim3(256), 0, s, nb, A, lda, info, info_base);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

template <typename T>
int launch_trsm_rlt_panel(int nb, int64_t nrows, const T* L, int64_t ldl, T* B, int64_t ldb,
                          hipStream_t s, bool l_transposed, bool unit) {
  if (nrows <= 0 || nb <= 0) return 0;
  if (nb > NB) {
    set_error("trsm block %d > %d", nb, NB);
    return DDLA_ERR_INVALID_ARG;
  }
  ProfScope prof(sizeof(T) == 8 ? "trsm_rlt_f64" : "trsm_rlt_f32", (double)nb * nb * nrows,
                 sizeof(T) * (2.0 * nrows * nb + 0.5 * nb * nb), s);
  if (l_transposed)
    hipLaunchKernelGGL((trsm_rlt_kernel<T, true>), dim3(ceil_div(nrows, 64)), dim3(256), 0, s, nb,
                       nrows, L, ldl, B, ldb, unit);
  else
    hipLaunchKernelGGL((trsm_rlt_kernel<T, false>), dim3(ceil_div(nrows, 64)), dim3(256), 0, s, nb,
                       nrows, L, ldl, B, ldb, unit);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

template <typename T>
int launch_getf2_nopiv(int nb, T* A, int64_t lda, int32_t* info, int info_base, hipStream_t s) {
  if (nb <= 0) return 0;
  if (nb > NB) {
    set_error("getf2 block %d > %d", nb, NB);
    return DDLA_ERR_INVALID_ARG;
  }
  ProfScope prof(sizeof(T) == 8 ? "getf2_nopiv_f64" : "getf2_nopiv_f32",
                 2.0 * nb * nb * nb / 3.0, 2.0 * sizeof(T) * nb * nb, s);
  hipLaunchKernelGGL(getf2_nopiv_kernel<T>, dim3(1), dim3(256), 0, s, nb, A, lda, info, info_base);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_dtranspose_inplace(int64_t n, double* a, int64_t lda, hipStream_t s) {
  if (n <= 0) return 0;
  const int nt = ceil_div(n, 32);
  hipLaunchKernelGGL(transpose_kernel, dim3(nt, nt), dim3(256), 0, s, n, a, lda);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ldlt_diag(int nb, double* A, int64_t lda, int32_t* info, int info_base,
                     hipStream_t s) {
  if (nb <= 0) return 0;
  if (nb > NB) {
    set_error("ldlt block %d > %d", nb, NB);
    return DDLA_ERR_INVALID_ARG;
  }
  ProfScope prof("ldlt_diag_f64", (double)nb * nb * nb / 3.0, 8.0 * nb * nb, s);
  hipLaunchKernelGGL(ldlt_diag_kernel, dim3(1), dim3(256), 0, s, nb, A, lda, info, info_base);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ldlt_scale(int64_t rows, int b, const double* A11, int64_t lda, double* A21,
                      double* Y, int64_t ldy, hipStream_t s) {
  const int64_t total = rows * b;
  if (total <= 0) return 0;
  hipLaunchKernelGGL(ldlt_scale_kernel, dim3(ceil_div(total, 256)), dim3(256), 0, s, rows, b, A11,
                     lda, A21, Y, ldy);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_scale_rows_by_inv_diag(int64_t n, int64_t nrhs, const double* A, int64_t lda,
                                  double* X, int64_t ldx, hipStream_t s) {
  const int64_t total = n * nrhs;
  if (total <= 0) return 0;
  hipLaunchKernelGGL(scale_rows_by_inv_diag_kernel, dim3(ceil_div(total, 256)), dim3(256), 0, s,
                     n, nrhs, A, lda, X, ldx);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

template int launch_potf2_ex<double>(int, double*, int64_t, int32_t*, int, hipStream_t);
template int launch_potf2_ex<float>(int, float*, int64_t, int32_t*, int, hipStream_t);
template int launch_trsm_rlt_panel<double>(int, int64_t, const double*, int64_t, double*, int64_t,
                                           hipStream_t, bool, bool);
template int launch_trsm_rlt_panel<float>(int, int64_t, const float*, int64_t, float*, int64_t,
                                          hipStream_t, bool, bool);
template int launch_getf2_nopiv<double>(int, double*, int64_t, int32_t*, int, hipStream_t);
template int launch_getf2_nopiv<float>(int, float*, int64_t, int32_t*, int, hipStream_t);

}  // namespace ddla
