// Triangular solves with few right-hand sides (the refinement loop's L and U solves).
//
// Replaces, for nrhs <= 8, the blocked launch-per-block TRSM behind PLUSolveInternel
// (DDLA/Factorizations/LU.cs:451-464) and CholeskySolveUncheck
// (DDLA/Factorizations/Cholesky.cs:210-226); the reference runs them through bli_dtrsm.
//
// One persistent cooperative launch solves op(T) X = B for all 64-row blocks:
//   * block row kb is owned by workgroup kb % G; every workgroup walks its blocks in
//     dependency order, so a block only waits on blocks owned by resident workgroups;
//   * for each earlier block kc it stages the 64 x 64 block of op(T) in LDS (coalesced
//     reads, transposed in LDS when op is a transpose), waits for X[kc] and accumulates
//     op(T)[kb,kc] X[kc];
//   * the diagonal block is applied as a product with its precomputed inverse
//     (trsv_dinv_kernel, once per factor), so the chain of dependent work per block is
//     a few small matrix-vector products;
//   * X[kc] is published as 64-bit words (32 bits of payload + 32-bit call tag, two words
//     per FP64 value) with single-copy-atomic agent-scope stores; a reader re-reads a word
//     until its tag is the current call's.  No fences, flags or counters are needed.
#include "ddla_internal.hpp"

namespace ddla {
namespace {

constexpr int BS = 64;    // block rows
constexpr int RMAX = 8;   // right-hand sides per launch
typedef unsigned long long u64;

template <typename T>
struct Words {
  static constexpr int N = sizeof(T) == 8 ? 2 : 1;
};

template <typename T>
__device__ __forceinline__ void publish(u64* p, T v, unsigned tag) {
  if constexpr (sizeof(T) == 8) {
    const u64 b = __builtin_bit_cast(u64, v);
    __hip_atomic_store(p, ((b >> 32) << 32) | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 1, (b << 32) | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    __hip_atomic_store(p, (static_cast<u64>(__builtin_bit_cast(unsigned, v)) << 32) | tag,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <typename T>
__device__ __forceinline__ T await(const u64* p, unsigned tag) {
  if constexpr (sizeof(T) == 8) {
    u64 a = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    u64 b = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (static_cast<unsigned>(a) != tag || static_cast<unsigned>(b) != tag) {
      __builtin_amdgcn_s_sleep(1);
      a = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      b = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return __builtin_bit_cast(T, ((a >> 32) << 32) | (b >> 32));
  } else {
    u64 a = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (static_cast<unsigned>(a) != tag) {
      __builtin_amdgcn_s_sleep(1);
      a = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return __builtin_bit_cast(T, static_cast<unsigned>(a >> 32));
  }
}

// element (i, k) of op(T) for a row-major T
template <typename T, bool TRANS>
__device__ __forceinline__ T opT(const T* A, int64_t lda, int64_t i, int64_t k) {
  return TRANS ? A[k * lda + i] : A[i * lda + k];
}

// Dinv[b] = inverse of the b-th 64 x 64 diagonal block of op(T) (row-major 64 x 64).
// One workgroup per block; lane j (< 64) solves op(T)_bb x = e_j by substitution.
template <typename T, bool LOWER, bool UNIT, bool TRANS>
__global__ __launch_bounds__(256) void trsv_dinv_kernel(int64_t n, const T* __restrict__ A,
                                                        int64_t lda, T* __restrict__ Dinv) {
  __shared__ T M[BS][BS + 1];
  __shared__ T X[BS][BS + 1];
  const int64_t b0 = (int64_t)blockIdx.x * BS;
  const int bs = (int)min<int64_t>(BS, n - b0);
  for (int idx = threadIdx.x; idx < BS * BS; idx += 256) {
    const int r = idx / BS, c = idx % BS;
    T v = T(0);
    if (r < bs && c < bs) v = TRANS ? A[(b0 + c) * lda + b0 + r] : A[(b0 + r) * lda + b0 + c];
    if (r >= bs || c >= bs) v = (r == c) ? T(1) : T(0);  // pad with the identity
    M[r][c] = v;
  }
  __syncthreads();
  const int j = threadIdx.x;
  if (j < BS) {
    if (LOWER) {
      for (int i = 0; i < BS; ++i) {
        T s = (i == j) ? T(1) : T(0);
        for (int k = j; k < i; ++k) s = fma(-M[i][k], X[k][j], s);
        X[i][j] = (UNIT || i >= bs) ? s : s / M[i][i];
      }
    } else {
      for (int i = BS - 1; i >= 0; --i) {
        T s = (i == j) ? T(1) : T(0);
        for (int k = i + 1; k <= j; ++k) s = fma(-M[i][k], X[k][j], s);
        X[i][j] = (UNIT || i >= bs) ? s : s / M[i][i];
      }
    }
  }
  __syncthreads();
  T* D = Dinv + (int64_t)blockIdx.x * BS * BS;
  for (int idx = threadIdx.x; idx < BS * BS; idx += 256) {
    const int r = idx / BS, c = idx % BS;
    const bool zero = LOWER ? (c > r) : (c < r);
    D[idx] = zero ? T(0) : X[r][c];
  }
}

template <typename T, bool LOWER, bool TRANS>
__global__ __launch_bounds__(256) void trsv_kernel(int64_t n, int nrhs, const T* __restrict__ A,
                                                   int64_t lda, const T* __restrict__ Dinv,
                                                   T* __restrict__ B, int64_t ldb,
                                                   u64* __restrict__ XW, unsigned tag) {
  __shared__ T Ts[BS][BS + 1];
  __shared__ T xs[BS][RMAX];
  __shared__ T red[4][BS][RMAX];
  const int nblk = (int)((n + BS - 1) / BS);
  const int G = gridDim.x;
  const int tid = threadIdx.x;
  const int row = tid & 63, q = tid >> 6;  // q: k-quarter of the block product
  constexpr int NW = Words<T>::N;

  for (int jb = blockIdx.x; jb < nblk; jb += G) {
    const int kb = LOWER ? jb : nblk - 1 - jb;  // dependency order
    const int64_t r0 = (int64_t)kb * BS;
    const int rows = (int)min<int64_t>(BS, n - r0);
    T part[RMAX];
#pragma unroll
    for (int c = 0; c < RMAX; ++c) part[c] = T(0);
    // off-diagonal blocks, in the order their X become available
    for (int jc = 0; jc < jb; ++jc) {
      const int kc = LOWER ? jc : nblk - 1 - jc;
      const int64_t c0 = (int64_t)kc * BS;
      const int cols = (int)min<int64_t>(BS, n - c0);
      __syncthreads();  // previous step finished with Ts / xs
      for (int idx = tid; idx < BS * BS; idx += 256) {
        int r, c;
        if (TRANS) {  // coalesced along T's rows = op(T)'s columns
          c = idx / BS;
          r = idx % BS;
        } else {
          r = idx / BS;
          c = idx % BS;
        }
        Ts[r][c] = (r < rows && c < cols) ? opT<T, TRANS>(A, lda, r0 + r, c0 + c) : T(0);
      }
      for (int idx = tid; idx < BS * nrhs; idx += 256) {
        const int k = idx / nrhs, c = idx % nrhs;
        xs[k][c] = (k < cols) ? await<T>(XW + ((c0 + k) * nrhs + c) * NW, tag) : T(0);
      }
      __syncthreads();
#pragma unroll 4
      for (int k = q * 16; k < q * 16 + 16; ++k) {
        const T a = Ts[row][k];
#pragma unroll
        for (int c = 0; c < RMAX; ++c)
          if (c < nrhs) part[c] = fma(a, xs[k][c], part[c]);
      }
    }
    // rhs = B[kb] - sum; then X[kb] = Dinv[kb] rhs
#pragma unroll
    for (int c = 0; c < RMAX; ++c)
      if (c < nrhs) red[q][row][c] = part[c];
    __syncthreads();
    for (int idx = tid; idx < BS * nrhs; idx += 256) {
      const int r = idx / nrhs, c = idx % nrhs;
      const T s = red[0][r][c] + red[1][r][c] + red[2][r][c] + red[3][r][c];
      xs[r][c] = (r < rows) ? B[(r0 + r) * ldb + c] - s : T(0);
    }
    const T* D = Dinv + (int64_t)kb * BS * BS;
    for (int idx = tid; idx < BS * BS; idx += 256) Ts[idx / BS][idx % BS] = D[idx];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < RMAX; ++c) part[c] = T(0);
#pragma unroll 4
    for (int k = q * 16; k < q * 16 + 16; ++k) {
      const T a = Ts[row][k];
#pragma unroll
      for (int c = 0; c < RMAX; ++c)
        if (c < nrhs) part[c] = fma(a, xs[k][c], part[c]);
    }
#pragma unroll
    for (int c = 0; c < RMAX; ++c)
      if (c < nrhs) red[q][row][c] = part[c];
    __syncthreads();
    for (int idx = tid; idx < BS * nrhs; idx += 256) {
      const int r = idx / nrhs, c = idx % nrhs;
      if (r < rows) {
        const T x = red[0][r][c] + red[1][r][c] + red[2][r][c] + red[3][r][c];
        publish<T>(XW + ((r0 + r) * nrhs + c) * NW, x, tag);
        B[(r0 + r) * ldb + c] = x;
      }
    }
  }
}

template <typename T>
struct TrsvState {
  unsigned tag = 0;
  void* ws = nullptr;
  size_t bytes = 0;
};

template <typename T, bool LOWER, bool TRANS>
int max_trsv_groups() {
  static int cached = -1;
  if (cached < 0) {
    int per_cu = 0;
    DDLA_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(trsv_kernel<T, LOWER, TRANS>), 256, 0));
    cached = per_cu * ctx().num_cus;
  }
  return cached;
}

}  // namespace

// Dinv (nblk x 64 x 64) for op(T); nblk = ceil(n/64)
template <typename T>
int trsv_prepare(bool lower, bool unit, bool trans, int64_t n, const T* A, int64_t lda, T* Dinv,
                 hipStream_t s) {
  const int nblk = ceil_div(n, BS);
#define TD(L, U, TR)                                                                            \
  hipLaunchKernelGGL((trsv_dinv_kernel<T, L, U, TR>), dim3(nblk), dim3(256), 0, s, n, A, lda, \
                     Dinv)
  // lower/upper refer to op(T)
  if (lower) {
    if (unit) { if (trans) TD(true, true, true); else TD(true, true, false); }
    else { if (trans) TD(true, false, true); else TD(true, false, false); }
  } else {
    if (unit) { if (trans) TD(false, true, true); else TD(false, true, false); }
    else { if (trans) TD(false, false, true); else TD(false, false, false); }
  }
#undef TD
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

size_t trsv_dinv_elems(int64_t n) { return (size_t)ceil_div(n, BS) * BS * BS; }

// op(T) X = B in place, nrhs <= 8, with Dinv from trsv_prepare (same lower/trans).
template <typename T>
int trsv_solve(bool lower, bool trans, int64_t n, int nrhs, const T* A, int64_t lda, const T* Dinv,
               T* B, int64_t ldb, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return 0;
  if (nrhs > RMAX) {
    set_error("trsv: nrhs %d > %d", nrhs, RMAX);
    return DDLA_ERR_INVALID_ARG;
  }
  static unsigned tag = 0;
  void* ws = nullptr;
  DDLA_RETURN_IF(workspace(WS_GMRES, (size_t)n * nrhs * 2 * sizeof(u64), &ws));
  static void* zeroed = nullptr;
  if (zeroed != ws) {
    DDLA_HIP_CHECK(hipMemsetAsync(ws, 0, (size_t)n * nrhs * 2 * sizeof(u64), s));
    zeroed = ws;
    tag = 0;
  }
  tag += 1;
  if (tag == 0) tag = 1;
  const int nblk = ceil_div(n, BS);
  ProfScope prof(sizeof(T) == 8 ? "trsv_f64" : "trsv_f32", (double)n * n * nrhs,
                 (0.5 * n * n + 2.0 * n * nrhs) * sizeof(T), s);
  u64* XW = static_cast<u64*>(ws);
  int64_t n_ = n;
  int nrhs_ = nrhs;
  int64_t lda_ = lda, ldb_ = ldb;
  unsigned tag_ = tag;
  void* args[] = {&n_, &nrhs_, &A, &lda_, &Dinv, &B, &ldb_, &XW, &tag_};
#define TS(L, TR)                                                                        \
  do {                                                                                   \
    const int G = std::min(nblk, max_trsv_groups<T, L, TR>());                           \
    DDLA_HIP_CHECK(hipLaunchCooperativeKernel(                                           \
        reinterpret_cast<const void*>(trsv_kernel<T, L, TR>), dim3(G), dim3(256), args, 0, s)); \
  } while (0)
  if (lower) {
    if (trans) TS(true, true); else TS(true, false);
  } else {
    if (trans) TS(false, true); else TS(false, false);
  }
#undef TS
  return 0;
}

template int trsv_prepare<float>(bool, bool, bool, int64_t, const float*, int64_t, float*, hipStream_t);
template int trsv_prepare<double>(bool, bool, bool, int64_t, const double*, int64_t, double*, hipStream_t);
template int trsv_solve<float>(bool, bool, int64_t, int, const float*, int64_t, const float*, float*,
                               int64_t, hipStream_t);
template int trsv_solve<double>(bool, bool, int64_t, int, const double*, int64_t, const double*,
                                double*, int64_t, hipStream_t);

}  // namespace ddla
