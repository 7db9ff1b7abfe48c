// Kernels of the blocked Cholesky (lower).  Reference: CholeskyLowerBlock /
// CholeskyLowerUnblock / Sqrt (DDLA/Factorizations/Cholesky.cs:127-160,228-236).
//
//   potf2_kernel      one workgroup factors the nb x nb diagonal block in LDS
//                     (sqrt, scale by the reciprocal like InvScaled, lower rank-1 update)
//   trsm_rlt_kernel   A21 <- A21 * L11^-T, one workgroup per 64 rows of A21, the lower
//                     triangle of L11 packed in LDS and the slab stored transposed so lanes
//                     run along rows
//   transpose_kernel  in-place transpose of a square matrix (the upper factorization is the
//                     lower one of the transposed storage)
#include "ddla_internal.hpp"

namespace ddla {
namespace {

constexpr int NB = 128;

template <typename T>
__global__ __launch_bounds__(256) void potf2_kernel(int nb, T* __restrict__ A, int64_t lda,
                                                    int32_t* __restrict__ info, int info_base) {
  __shared__ T t[NB * (NB + 1)];
  __shared__ T s_rinv;
  __shared__ int s_stop;
  constexpr int LD = NB + 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
  for (int idx = tid; idx < nb * nb; idx += 256) {
    const int r = idx / nb, c = idx % nb;
    if (c <= r) t[r * LD + c] = A[(int64_t)r * lda + c];
  }
  if (tid == 0) s_stop = 0;
  __syncthreads();
  for (int i = 0; i < nb; ++i) {
    if (tid == 0) {
      const T a = t[i * LD + i];
      if (!(a > static_cast<T>(0))) {
        s_stop = 1;
        atomicCAS(info, 0, info_base + i + 1);
      } else {
        const T d = sqrt(a);
        t[i * LD + i] = d;
        s_rinv = static_cast<T>(1) / d;
      }
    }
    __syncthreads();
    if (s_stop) break;
    const T rinv = s_rinv;
    for (int r = i + 1 + tid; r < nb; r += 256) t[r * LD + i] *= rinv;
    __syncthreads();
    // lower rank-1 update; lane -> columns c0, c1, wave -> rows wave + 4k, batches of rows
    // whose LDS loads are issued before their stores
    const int c0 = i + 1 + lane, c1 = c0 + 64;
    const T w0 = c0 < nb ? t[c0 * LD + i] : T(0), w1 = c1 < nb ? t[c1 * LD + i] : T(0);
    constexpr int CH = 8;
    for (int k0 = 0; k0 < NB / 4; k0 += CH) {
      T x[CH], a0[CH], a1[CH];
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const T* tr = t + (wave + 4 * (k0 + k)) * LD;
        x[k] = tr[i];
        a0[k] = tr[c0 < NB ? c0 : 0];
        a1[k] = tr[c1 < NB ? c1 : 0];
      }
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int r = wave + 4 * (k0 + k);
        if (r > i && r < nb) {
          T* tr = t + r * LD;
          if (c0 <= r) tr[c0] = fma(w0, -x[k], a0[k]);
          if (c1 <= r) tr[c1] = fma(w1, -x[k], a1[k]);
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
  for (int idx = tid; idx < nb * nb; idx += 256) {
    const int r = idx / nb, c = idx % nb;
    t[r * LD + c] = A[(int64_t)r * lda + c];
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

// B (nrows x nb) <- B * L^-T, L lower nb x nb (non-unit).  Row b of B solves L x^T = b^T.
// TL: L is given by its transpose, i.e. L[r][c] is read from U[c][r] (B <- B * U^-1 for an
// upper U stored at the same address).
template <typename T, bool TL>
__global__ __launch_bounds__(256) void trsm_rlt_kernel(int nb, int64_t nrows,
                                                       const T* __restrict__ L, int64_t ldl,
                                                       T* __restrict__ B, int64_t ldb) {
  __shared__ T Ls[NB * (NB + 1) / 2];
  constexpr int SL = 65;   // padded row of the transposed slab
  __shared__ T S[NB * SL];  // transposed slab: S[c*SL + r]
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  const int rh = (int)min<int64_t>(64, nrows - r0);
  const int npk = nb * (nb + 1) / 2;
  for (int p = tid; p < npk; p += 256) {  // flat packed index: independent loads
    int r = (int)((sqrtf(8.0f * p + 1.0f) - 1.0f) * 0.5f);
    while (r * (r + 1) / 2 > p) --r;
    while ((r + 1) * (r + 2) / 2 <= p) ++r;
    const int c = p - r * (r + 1) / 2;
    Ls[p] = TL ? L[(int64_t)c * ldl + r] : L[(int64_t)r * ldl + c];
  }
  for (int idx = tid; idx < 64 * nb; idx += 256) {
    const int r = idx / nb, c = idx % nb;
    S[c * SL + r] = (r < rh) ? B[(r0 + r) * ldb + c] : static_cast<T>(0);
  }
  __syncthreads();
  const int rr = tid & 63, g = tid >> 6;
  if (g == 0) S[rr] = S[rr] / Ls[0];
  __syncthreads();
  // column c of the slab (c = g + 4k) is owned by group g; each step loads a batch of
  // owned columns before storing them (one LDS round trip per batch)
  for (int i = 0; i < nb - 1; ++i) {
    const T x = S[i * SL + rr];
    constexpr int CH = 8;
    for (int k0 = 0; k0 < NB / 4; k0 += CH) {
      T v[CH], l[CH];
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int c = g + 4 * (k0 + k);
        v[k] = S[c * SL + rr];
        l[k] = (c > i && c < nb) ? Ls[c * (c + 1) / 2 + i] : T(0);
      }
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int c = g + 4 * (k0 + k);
        if (c > i && c < nb) {
          T w = fma(-l[k], x, v[k]);
          if (c == i + 1) w = w / Ls[c * (c + 1) / 2 + c];
          S[c * SL + rr] = w;
        }
      }
    }
    __syncthreads();
  }
  for (int idx = tid; idx < 64 * nb; idx += 256) {
    const int r = idx / nb, c = idx % nb;
    if (r < rh) B[(r0 + r) * ldb + c] = S[c * SL + r];
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

}  // namespace

template <typename T>
int launch_potf2_ex(int nb, T* A, int64_t lda, int32_t* info, int info_base, hipStream_t s) {
  if (nb <= 0) return 0;
  if (nb > NB) {
    set_error("potf2 block %d > %d", nb, NB);
    return DDLA_ERR_INVALID_ARG;
  }
  ProfScope prof(sizeof(T) == 8 ? "potf2_f64" : "potf2_f32", (double)nb * nb * nb / 3.0,
                 sizeof(T) * (double)nb * nb, s);
  hipLaunchKernelGGL(potf2_kernel<T>, dim3(1), dim3(256), 0, s, nb, A, lda, info, info_base);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

template <typename T>
int launch_trsm_rlt_panel(int nb, int64_t nrows, const T* L, int64_t ldl, T* B, int64_t ldb,
                          hipStream_t s, bool l_transposed) {
  if (nrows <= 0 || nb <= 0) return 0;
  if (nb > NB) {
    set_error("trsm block %d > %d", nb, NB);
    return DDLA_ERR_INVALID_ARG;
  }
  ProfScope prof(sizeof(T) == 8 ? "trsm_rlt_f64" : "trsm_rlt_f32", (double)nb * nb * nrows,
                 sizeof(T) * (2.0 * nrows * nb + 0.5 * nb * nb), s);
  if (l_transposed)
    hipLaunchKernelGGL((trsm_rlt_kernel<T, true>), dim3(ceil_div(nrows, 64)), dim3(256), 0, s, nb,
                       nrows, L, ldl, B, ldb);
  else
    hipLaunchKernelGGL((trsm_rlt_kernel<T, false>), dim3(ceil_div(nrows, 64)), dim3(256), 0, s, nb,
                       nrows, L, ldl, B, ldb);
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

template int launch_potf2_ex<double>(int, double*, int64_t, int32_t*, int, hipStream_t);
template int launch_potf2_ex<float>(int, float*, int64_t, int32_t*, int, hipStream_t);
template int launch_trsm_rlt_panel<double>(int, int64_t, const double*, int64_t, double*, int64_t,
                                           hipStream_t, bool);
template int launch_trsm_rlt_panel<float>(int, int64_t, const float*, int64_t, float*, int64_t,
                                          hipStream_t, bool);
template int launch_getf2_nopiv<double>(int, double*, int64_t, int32_t*, int, hipStream_t);
template int launch_getf2_nopiv<float>(int, float*, int64_t, int32_t*, int, hipStream_t);

}  // namespace ddla
