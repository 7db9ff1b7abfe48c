// Mixed-precision LU and Cholesky with iterative refinement to FP64 accuracy.
//
// No counterpart in the reference (DDLA is FP64-only).  The factorizations follow the same
// blocked algorithms as the FP64 path (DDLA/Factorizations/LU.cs:289-342,
// DDLA/Factorizations/Cholesky.cs:142-160) on an FP32 working copy:
//   * panel / diagonal block and TRSM in FP32 (partial pivoting, same kernels as FP64);
//   * the trailing update C <- C - A B runs on the FP16 or BF16 matrix cores
//     (v_mfma_f32_16x16x32_{f16,bf16}) with FP32 accumulation; operands are rounded
//     while they are staged from HBM into LDS, after multiplication by a per-128-block
//     power-of-two scale (max |x| of the block mapped to (2^13, 2^14], so FP16 neither
//     overflows nor loses small entries to subnormals); the epilogue divides the scales
//     out exactly;
//   * refinement: X += (LU)^-1 P R with R = B - A X in FP64 (GEMM on the FP64 matrix
//     cores), stopping when max_j ||r_j||_inf / (||A||_inf ||x_j||_inf) <= tol.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ddla_internal.hpp"

namespace ddla {

int lu_bookkeeping(int64_t n, int32_t** dinfo, int** counts, int** lists);

namespace {

constexpr int NBLK = 128;
constexpr int HB = 128;  // scale block (rows of A / columns of B) = GEMM tile

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <bool BF16>
struct LowP;
template <>
struct LowP<false> {
  typedef half8 v8;
  typedef _Float16 e;
  static __device__ __forceinline__ f4 mfma(v8 a, v8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct LowP<true> {
  typedef bf16x8 v8;
  typedef __bf16 e;
  static __device__ __forceinline__ f4 mfma(v8 a, v8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

// power-of-two scale of a block: 1 for bf16 (FP32 exponent range) or an empty block
__device__ __forceinline__ float block_scale(float amax, bool bf16) {
  if (bf16 || !(amax > 0.0f) || !isfinite(amax)) return 1.0f;
  const int e = (int)ceilf(log2f(amax));
  return exp2f((float)(14 - e));
}

// Operand preparation for the trailing update: rows of an operand (rows of A, columns of B)
// in blocks of 128 share one power-of-two scale; each scaled FP32 value x is split into
// TERMS low-precision parts x = h + m + l with h = rn(x), m = rn(x - h), l = rn(x - h - m),
// written K-contiguous with rows padded to Kp (a multiple of 32) with zeros:
// out[t][r * Kp + k].  One workgroup per 128-row block (KC: x(r,k) = src[r ld + k]; else
// src[k ld + r]).  Every thread issues all of its sixteen 4-wide reads before it uses one
// (the block is latency-bound, not bandwidth-bound: no read may wait for the previous one); the transposed form goes through LDS so
// that both sides stay coalesced.  Done once per update, not once per output tile.
template <bool BF16, int TERMS, bool KC>
__global__ __launch_bounds__(256) void split_kernel(int rows, int K, int Kp,
                                                    const float* __restrict__ src, int64_t ld,
                                                    typename LowP<BF16>::e* __restrict__ out,
                                                    float* __restrict__ scales) {
  typedef typename LowP<BF16>::e E;
  typedef E e4 __attribute__((ext_vector_type(4)));
  __shared__ float xs[KC ? 1 : HB][HB + 1];
  __shared__ float red[4];
  const int r0 = blockIdx.x * HB;
  const int nr = min(HB, rows - r0);
  const int tid = threadIdx.x;
  const bool vec = (ld & 3) == 0 && (reinterpret_cast<uintptr_t>(src) & 15) == 0;
  // item f = i 256 + tid covers (a, b4 .. b4 + 3), a = f / 32, b4 = 4 (f mod 32): for KC row a
  // of the block and k = b4.., else k = a and rows b4.. of the block
  f4 v[16];
  if (vec) {
    // whole 4-wide items: branch-free 16-byte reads (an item outside reads src itself and is
    // zeroed), all in flight together; items cut by the edge are read element-wise after
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int f = i * 256 + tid, a = f >> 5, b4 = (f & 31) * 4;
      const bool full = (KC ? a < nr : a < K) && b4 + 3 < (KC ? K : nr);
      const float* p = src + (int64_t)(KC ? r0 + a : a) * ld + (KC ? b4 : r0 + b4);
      v[i] = *reinterpret_cast<const f4*>(full ? p : src);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int f = i * 256 + tid, a = f >> 5, b4 = (f & 31) * 4;
      const bool aok = KC ? a < nr : a < K;
      const int lim = KC ? K : nr;
      if (!(aok && b4 + 3 < lim)) {
        const float* p = src + (int64_t)(KC ? r0 + a : a) * ld + (KC ? b4 : r0 + b4);
        f4 x = {0.f, 0.f, 0.f, 0.f};
        if (aok) {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (b4 + e < lim) x[e] = p[e];
        }
        v[i] = x;
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int f = i * 256 + tid, a = f >> 5, b4 = (f & 31) * 4;
      const bool aok = KC ? a < nr : a < K;
      const int lim = KC ? K : nr;
      const int64_t off = (int64_t)(KC ? r0 + a : a) * ld + (KC ? b4 : r0 + b4);
#pragma unroll
      for (int e = 0; e < 4; ++e) v[i][e] = src[(aok && b4 + e < lim) ? off + e : 0];
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int f = i * 256 + tid, a = f >> 5, b4 = (f & 31) * 4;
      const bool aok = KC ? a < nr : a < K;
      const int lim = KC ? K : nr;
#pragma unroll
      for (int e = 0; e < 4; ++e) v[i][e] = (aok && b4 + e < lim) ? v[i][e] : 0.f;
    }
  }
  float mx = 0.0f;
#pragma unroll
  for (int i = 0; i < 16; ++i)
    mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v[i][0]), fabsf(v[i][1])), fmaxf(fabsf(v[i][2]), fabsf(v[i][3]))));
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
  if ((tid & 63) == 0) red[tid >> 6] = mx;
  if (!KC) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int f = i * 256 + tid, a = f >> 5, b4 = (f & 31) * 4;
#pragma unroll
      for (int e = 0; e < 4; ++e) xs[b4 + e][a] = v[i][e];
    }
  }
  __syncthreads();
  const float sc = block_scale(fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3])), BF16);
  if (tid == 0) scales[blockIdx.x] = sc;
  const size_t plane = (size_t)rows * Kp;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int f = i * 256 + tid, a = f >> 5, b4 = (f & 31) * 4;  // row a, k = b4 ..
    if (a >= nr || b4 >= Kp) continue;
    f4 x;
    if (KC) {
      x = v[i];
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) x[e] = xs[a][b4 + e];
    }
    x *= sc;
    E* o = out + (size_t)(r0 + a) * Kp + b4;
#pragma unroll
    for (int t = 0; t < TERMS; ++t) {
      e4 h;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        h[e] = static_cast<E>(x[e]);
        x[e] = x[e] - static_cast<float>(h[e]);
      }
      *reinterpret_cast<e4*>(o + t * plane) = h;
    }
  }
}

// C (FP32, m x n) += alpha * A B from the split operands Ap[t pa + i K + k], Bp[t pb + j K + k]
// (both K-contiguous, K a multiple of 32 with zero padding) with FP32 accumulation on the matrix cores (v_mfma_f32_16x16x32_{f16,bf16}).
// The tile product sums the part products whose combined weight reaches the FP32 level
// (TERMS=1: hh; 2: hh+hm+mh; 3: hh+hm+mh+hl+lh+mm), small ones first.  128 x 128 tile per
// workgroup, K step 32 through LDS with register prefetch; the C tile is read at the start
// (its latency overlaps the products) and updated once at the end.  TRI == 1: lower triangle.
template <bool BF16, int TERMS, int TRI>
__global__ __launch_bounds__(256) void hgemm_kernel(int M, int N, int K, float alpha,
                                                    const typename LowP<BF16>::e* __restrict__ Ap,
                                                    const typename LowP<BF16>::e* __restrict__ Bp,
                                                    float* __restrict__ C, int64_t ldc,
                                                    const float* __restrict__ sa,
                                                    const float* __restrict__ sb, int tiles_m,
                                                    int tiles_n, int64_t pa, int64_t pb) {
  typedef LowP<BF16> LP;
  typedef typename LP::v8 v8;
  typedef typename LP::e E;
  constexpr int BM = 128, BN = 128, BK = 32, PK = BK + 8;  // padded rows: conflict-free b128 reads
  __shared__ __attribute__((aligned(16))) E As[TERMS][BM * PK];
  __shared__ __attribute__((aligned(16))) E Bs[TERMS][BN * PK];

  int tmi, tni;
  if (TRI == 1)  // launched over the lower tiles only
    lower_tile_of(blockIdx.x, tiles_m, tiles_n, tmi, tni);
  else
    gemm_tile_of(blockIdx.x, tiles_m, tiles_n, tmi, tni);
  const int m0 = tmi * BM, n0 = tni * BN;
  if (TRI == 1 && n0 > m0 + BM - 1) return;
  const float scA = sa[tmi], scB = sb[tni];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int l16 = lane & 15, lk = lane >> 4;

  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  // staging: per term, 2 items of 8 k-values per lane for A and for B
  v8 ra[TERMS][2], rb[TERMS][2];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int v = tid + 256 * q;
      const int row = v >> 2, kg = (v & 3) * 8;
      const int gi = m0 + row, gj = n0 + row, gk = k0 + kg;
#pragma unroll
      for (int t = 0; t < TERMS; ++t) {
        v8 za = {}, zb = {};
        ra[t][q] = gi < M ? *reinterpret_cast<const v8*>(Ap + t * pa + (size_t)gi * K + gk) : za;
        rb[t][q] = gj < N ? *reinterpret_cast<const v8*>(Bp + t * pb + (size_t)gj * K + gk) : zb;
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int v = tid + 256 * q;
      const int row = v >> 2, kg = (v & 3) * 8;
#pragma unroll
      for (int t = 0; t < TERMS; ++t) {
        *reinterpret_cast<v8*>(&As[t][row * PK + kg]) = ra[t][q];
        *reinterpret_cast<v8*>(&Bs[t][row * PK + kg]) = rb[t][q];
      }
    }
  };

  const int nk = K / BK;
  load(0);
  // The products run as B^T A^T (operands swapped), so each lane's four results are four
  // consecutive columns of one row of C: 16-byte C accesses.  C is read right after the first
  // operands (its latency overlaps the products; requested after them, it does not delay them).
  const bool vec = (ldc & 3) == 0 && (reinterpret_cast<uintptr_t>(C) & 15) == 0;
  f4 cv[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gi = m0 + wm * 64 + i * 16 + l16;
      const int gj = n0 + wn * 64 + j * 16 + lk * 4;
      const float* p = C + (int64_t)gi * ldc + gj;
      if (vec && gi < M && gj + 3 < N) {
        cv[i][j] = *reinterpret_cast<const f4*>(p);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) cv[i][j][r] = (gi < M && gj + r < N) ? p[r] : 0.f;
      }
    }
  for (int kt = 0; kt < nk; ++kt) {
    __syncthreads();  // the previous step's reads of As / Bs are done
    store();
    __syncthreads();
    if (kt + 1 < nk) load((kt + 1) * BK);
    v8 a[TERMS][4], b[TERMS][4];
#pragma unroll
    for (int t = 0; t < TERMS; ++t) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        a[t][i] = *reinterpret_cast<const v8*>(&As[t][(wm * 64 + i * 16 + l16) * PK + lk * 8]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        b[t][j] = *reinterpret_cast<const v8*>(&Bs[t][(wn * 64 + j * 16 + l16) * PK + lk * 8]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (TERMS >= 3) {
          acc[i][j] = LP::mfma(b[1][j], a[1][i], acc[i][j]);
          acc[i][j] = LP::mfma(b[0][j], a[2][i], acc[i][j]);
          acc[i][j] = LP::mfma(b[2][j], a[0][i], acc[i][j]);
        }
        if (TERMS >= 2) {
          acc[i][j] = LP::mfma(b[0][j], a[1][i], acc[i][j]);
          acc[i][j] = LP::mfma(b[1][j], a[0][i], acc[i][j]);
        }
        acc[i][j] = LP::mfma(b[0][j], a[0][i], acc[i][j]);
      }
  }
  const float f = alpha / (scA * scB);  // exact: the scales are powers of two
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gi = m0 + wm * 64 + i * 16 + l16;
      const int gj = n0 + wn * 64 + j * 16 + lk * 4;
      float* p = C + (int64_t)gi * ldc + gj;
      f4 o;
#pragma unroll
      for (int r = 0; r < 4; ++r) o[r] = fmaf(f, acc[i][j][r], cv[i][j][r]);
      if (vec && gi < M && gj + 3 < N && (TRI == 0 || gj + 3 <= gi)) {
        *reinterpret_cast<f4*>(p) = o;
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (gi < M && gj + r < N && (TRI == 0 || gj + r <= gi)) p[r] = o[r];
      }
    }
}

__global__ void d2s_kernel(int64_t m, int64_t n, const double* __restrict__ src, int64_t lds,
                           float* __restrict__ dst, int64_t ldd) {
  const int64_t total = m * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / n, j = t % n;
    dst[i * ldd + j] = static_cast<float>(src[i * lds + j]);
  }
}

// Kernels of a refinement step take `skip`, the done word of the device-side convergence
// test (IrState::done, or null): once it is set, the steps queued after convergence do nothing.
__device__ __forceinline__ bool skipped(const int* skip) { return skip && *skip; }

// dst (float, n x nrhs contiguous) = (float) src[perm[i]]  (perm == nullptr: identity)
__global__ void gather_d2s_kernel(int64_t n, int64_t nrhs, const double* __restrict__ src,
                                  int64_t lds, const int* __restrict__ perm,
                                  float* __restrict__ dst, const int* __restrict__ skip) {
  if (skipped(skip)) return;
  const int64_t total = n * nrhs;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / nrhs, j = t % nrhs;
    const int64_t si = perm ? perm[i] : i;
    dst[t] = static_cast<float>(src[si * lds + j]);
  }
}

// X += (double) D ; R = B
__global__ void update_x_kernel(int64_t n, int64_t nrhs, const float* __restrict__ D,
                                double* __restrict__ X, int64_t ldx, const double* __restrict__ B,
                                int64_t ldb, double* __restrict__ R, const int* __restrict__ skip) {
  if (skipped(skip)) return;
  const int64_t total = n * nrhs;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / nrhs, j = t % nrhs;
    X[i * ldx + j] += static_cast<double>(D[t]);
    R[t] = B[i * ldb + j];
  }
}

// out[i] = sum_j |A[i][j]| (one wave per row)
__global__ __launch_bounds__(256) void row_abs_sum_kernel(int64_t n, const double* __restrict__ A,
                                                          int64_t lda, double* __restrict__ out) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double s = 0.0;
  const double* a = A + row * lda;
  for (int64_t j = lane; j < n; j += 64) s += fabs(a[j]);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) out[row] = s;
}

// dst = (float) A (as d2s_kernel) and out[i] = sum_j |A[i][j]| (as row_abs_sum_kernel, same
// order of the sums) in one pass over A: one wave per row
__global__ __launch_bounds__(256) void d2s_rowsum_kernel(int64_t n, const double* __restrict__ A,
                                                         int64_t lda, float* __restrict__ dst,
                                                         int64_t ldd, double* __restrict__ out) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double s = 0.0;
  const double* a = A + row * lda;
  float* d = dst + row * ldd;
#pragma unroll 4
  for (int64_t j = lane; j < n; j += 64) {
    const double v = a[j];
    d[j] = static_cast<float>(v);
    s += fabs(v);
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) out[row] = s;
}

// out[0] = max_i v[i] over n values (single workgroup)
__global__ __launch_bounds__(1024) void max_kernel(int64_t n, const double* __restrict__ v,
                                                   double* __restrict__ out) {
  __shared__ double red[16];
  double m = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) m = fmax(m, v[i]);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 16; ++w) m = fmax(m, red[w]);
    *out = m;
  }
}

// R[i][c] -= sum_k A[i][k] X[k][c] for NR right-hand sides: one wave per RPW rows of the
// FP64 matrix, lanes along the rows (coalesced 512-byte reads; each X value serves RPW rows),
// a butterfly sum at the end.  The refinement's residual reads A once per step, so this runs
// at HBM speed.
template <int NR>
__global__ __launch_bounds__(256) void dgemv_resid_kernel(int64_t n, const double* __restrict__ A,
                                                          int64_t lda, const double* __restrict__ X,
                                                          int64_t ldx, double* __restrict__ R,
                                                          int64_t ldr, int nrhs,
                                                          const int* __restrict__ skip) {
  if (skipped(skip)) return;
  constexpr int RPW = 4;
  const int64_t row0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * RPW;
  if (row0 >= n) return;
  const int lane = threadIdx.x & 63;
  const double* a[RPW];
#pragma unroll
  for (int q = 0; q < RPW; ++q) a[q] = A + min(row0 + q, n - 1) * lda;
  double acc[RPW][NR];
#pragma unroll
  for (int q = 0; q < RPW; ++q)
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[q][c] = 0.0;
#pragma unroll 2
  for (int64_t k = lane; k < n; k += 64) {
    double xv[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) xv[c] = c < nrhs ? X[k * ldx + c] : 0.0;
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      const double av = a[q][k];
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[q][c] = fma(av, xv[c], acc[q][c]);
    }
  }
#pragma unroll
  for (int q = 0; q < RPW; ++q)
#pragma unroll
    for (int c = 0; c < NR; ++c)
      for (int off = 32; off > 0; off >>= 1) acc[q][c] += __shfl_xor(acc[q][c], off);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < RPW; ++q)
#pragma unroll
      for (int c = 0; c < NR; ++c)
        if (c < nrhs && row0 + q < n) R[(row0 + q) * ldr + c] -= acc[q][c];
  }
}

__global__ void zero_u64_kernel(int64_t n, unsigned long long* p, const int* __restrict__ skip) {
  if (skipped(skip)) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// Device-side state of a refinement: the convergence test runs on the GPU after every step,
// so the host queues steps without waiting on each one.
struct IrState {
  int done;     // 1: converged, stagnated, not finite or out of steps (later steps do nothing)
  int iters;    // steps taken
  int slow;     // consecutive steps that lowered the backward error by less than 10%
  unsigned arrived;  // workgroups of the current step's norm launch that have finished
  double berr;  // backward error after the last step
  double prev;  // backward error before it
};

__global__ void ir_state_init_kernel(IrState* st, unsigned long long* norms, int nwords) {
  for (int i = 0; i < nwords; ++i) norms[i] = 0ull;
  st->done = 0;
  st->iters = 0;
  st->slow = 0;
  st->arrived = 0;
  st->berr = INFINITY;
  st->prev = INFINITY;
}

// The convergence test of one refinement step, in FP64 on the GPU: berr = max_j ||r_j||_inf /
// (||A||_inf ||x_j||_inf) from the norm words (see ir_norms_check_kernel); done when
// berr <= tol, after max_iters steps, when the error is not finite, or when three steps in a
// row lower it by less than 10% (stagnation).
__device__ void ir_check(int nrhs, const unsigned long long* norms, double na, double tol,
                         int max_iters, IrState* st) {
  double worst = 0.0;
  for (int j = 0; j < nrhs; ++j) {
    const double r = __longlong_as_double((long long)__hip_atomic_load(
        norms + 2 * j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    const double x = __longlong_as_double((long long)__hip_atomic_load(
        norms + 2 * j + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    double e = (x > 0.0 && na > 0.0) ? r / (na * x) : (r == 0.0 ? 0.0 : INFINITY);
    if (!isfinite(r) || !isfinite(x) || isnan(e)) e = INFINITY;  // NaN / Inf: failed
    worst = fmax(worst, e);
  }
  st->iters += 1;
  st->berr = worst;
  if (!(worst > tol) || !isfinite(worst) || st->iters >= max_iters) {
    st->done = 1;
    return;
  }
  st->slow = (worst > 0.9 * st->prev) ? st->slow + 1 : 0;
  if (st->slow >= 3) st->done = 1;
  st->prev = worst;
}

// One launch per refinement step for the norms and the test: norms[2j] = max_i |R[i][j]|,
// norms[2j+1] = max_i |X[i][j]| (bit patterns of non-negative doubles, NaN as infinity;
// atomic max, so the words do not depend on the order in which workgroups arrive), column
// j = blockIdx.y.  The last workgroup to arrive runs ir_check, then clears the words and the
// arrival count for the next step (ir_state_init_kernel clears them before the first).  A
// finished refinement (st->done, set only by a last workgroup, after every workgroup of its
// launch has read it) makes the launch a no-op.
__global__ __launch_bounds__(256) void ir_norms_check_kernel(
    int64_t n, int64_t nrhs, const double* __restrict__ R, const double* __restrict__ X,
    int64_t ldx, unsigned long long* __restrict__ norms, const double* __restrict__ normA,
    double tol, int max_iters, IrState* __restrict__ st) {
  if (st->done) return;
  const int64_t j = blockIdx.y;
  double rm = 0.0, xm = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double r = fabs(R[i * nrhs + j]), x = fabs(X[i * ldx + j]);
    if (r != r) r = INFINITY;
    if (x != x) x = INFINITY;
    rm = fmax(rm, r);
    xm = fmax(xm, x);
  }
  for (int off = 32; off > 0; off >>= 1) {
    rm = fmax(rm, __shfl_xor(rm, off));
    xm = fmax(xm, __shfl_xor(xm, off));
  }
  if ((threadIdx.x & 63) == 0) {
    __hip_atomic_fetch_max(norms + 2 * j, (unsigned long long)__double_as_longlong(rm),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_max(norms + 2 * j + 1, (unsigned long long)__double_as_longlong(xm),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();  // this wave's maxima are performed before the barrier
  }
  __shared__ int last;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned total = gridDim.x * gridDim.y;
    last = __hip_atomic_fetch_add(&st->arrived, 1u, __ATOMIC_ACQ_REL,
                                  __HIP_MEMORY_SCOPE_AGENT) == total - 1;
    if (last) {
      ir_check((int)nrhs, norms, *normA, tol, max_iters, st);
      for (int w = 0; w < 2 * (int)nrhs; ++w)
        __hip_atomic_store(norms + w, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      st->arrived = 0;
    }
  }
}

// out[2j] = max_i |X[i][j] - Xr[i][j]|, out[2j+1] = max_i |Xr[i][j]| (bit patterns of
// non-negative doubles, NaN as infinity; out cleared first), column j = blockIdx.y
__global__ __launch_bounds__(256) void fwd_err_norms_kernel(int64_t n, const double* __restrict__ X,
                                                            int64_t ldx,
                                                            const double* __restrict__ Xr,
                                                            int64_t ldr,
                                                            unsigned long long* __restrict__ out) {
  const int64_t j = blockIdx.y;
  double dm = 0.0, rm = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double xr = Xr[i * ldr + j];
    double d = fabs(X[i * ldx + j] - xr), r = fabs(xr);
    if (d != d) d = INFINITY;
    if (r != r) r = INFINITY;
    dm = fmax(dm, d);
    rm = fmax(rm, r);
  }
  for (int off = 32; off > 0; off >>= 1) {
    dm = fmax(dm, __shfl_xor(dm, off));
    rm = fmax(rm, __shfl_xor(rm, off));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(out + 2 * j, __double_as_longlong(dm));
    atomicMax(out + 2 * j + 1, __double_as_longlong(rm));
  }
}

int grid_for(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return static_cast<int>(g);
}

// ------------------------------------------------ random butterfly transforms (depth 2) ---
// A_r = U^T A V with U = diag(B1, B2) Bn: Bn is a size-np butterfly (1/sqrt2)[R0 R1; R0 -R1],
// B1 and B2 size-np/2 butterflies; the diagonals are exp(u/10), u uniform in [-1/2, 1/2)
// (Parker's recursive butterflies as used for pivot-free LU by Baboulin, Dongarra,
// Herrmann and Tomov, ACM TOMS 39(2), 2013).  Index groups g(i0) = {i0, i0+np/4, i0+np/2,
// i0+3np/4} are closed under U, so every transform is a 4 x 4 matrix per group.
// R layout per transform: a[np/2] b[np/2] c[np/4] d[np/4] e[np/4] f[np/4].
__global__ void rbt_diag_kernel(int64_t count, uint64_t seed, float* __restrict__ R) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count;
       t += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + static_cast<uint64_t>(t + 1) * 0x9E3779B97F4A7C15ull;
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = static_cast<double>(z >> 11) * 0x1.0p-53 - 0.5;
    R[t] = static_cast<float>(exp(u / 10.0));
  }
}

// Ug: rows/columns ordered as the group g(i0)
__device__ __forceinline__ void rbt_group(const float* __restrict__ R, int64_t np, int64_t i0,
                                          double M[4][4]) {
  const int64_t n2 = np / 2, n4 = np / 4;
  const double a0 = R[i0], a1 = R[i0 + n4], b0 = R[n2 + i0], b1 = R[n2 + i0 + n4];
  const double c = R[np + i0], d = R[np + n4 + i0], e = R[np + 2 * n4 + i0],
               f = R[np + 3 * n4 + i0];
  M[0][0] = 0.5 * c * a0; M[0][1] = 0.5 * d * a1;  M[0][2] = 0.5 * c * b0;  M[0][3] = 0.5 * d * b1;
  M[1][0] = 0.5 * c * a0; M[1][1] = -0.5 * d * a1; M[1][2] = 0.5 * c * b0;  M[1][3] = -0.5 * d * b1;
  M[2][0] = 0.5 * e * a0; M[2][1] = 0.5 * f * a1;  M[2][2] = -0.5 * e * b0; M[2][3] = -0.5 * f * b1;
  M[3][0] = 0.5 * e * a0; M[3][1] = -0.5 * f * a1; M[3][2] = -0.5 * e * b0; M[3][3] = 0.5 * f * b1;
}

// W (np x np FP32) = U^T [A 0; 0 I] V, A n x n FP64
__global__ __launch_bounds__(256) void rbt_matrix_kernel(int64_t n, int64_t np,
                                                         const double* __restrict__ A, int64_t lda,
                                                         const float* __restrict__ RU,
                                                         const float* __restrict__ RV,
                                                         float* __restrict__ W) {
  const int64_t n4 = np / 4;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n4 * n4;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i0 = t / n4, j0 = t % n4;
    double U[4][4], V[4][4], a[4][4], t1[4][4];
    rbt_group(RU, np, i0, U);
    rbt_group(RV, np, j0, V);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t gi = i0 + k * n4, gj = j0 + q * n4;
        a[k][q] = (gi < n && gj < n) ? A[gi * lda + gj] : (gi == gj ? 1.0 : 0.0);
      }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        t1[p][q] = U[0][p] * a[0][q] + U[1][p] * a[1][q] + U[2][p] * a[2][q] + U[3][p] * a[3][q];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double v = t1[p][0] * V[0][q] + t1[p][1] * V[1][q] + t1[p][2] * V[2][q] +
                         t1[p][3] * V[3][q];
        W[(i0 + p * n4) * np + j0 + q * n4] = static_cast<float>(v);
      }
  }
}

// D (np x nrhs FP32) = U^T [R; 0], R n x nrhs FP64 (contiguous)
__global__ void rbt_vec_kernel(int64_t n, int64_t np, int64_t nrhs, const double* __restrict__ Rv,
                               const float* __restrict__ RU, float* __restrict__ D,
                               const int* __restrict__ skip) {
  if (skipped(skip)) return;
  const int64_t n4 = np / 4;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n4 * nrhs;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i0 = t / nrhs, c = t % nrhs;
    double U[4][4], r[4];
    rbt_group(RU, np, i0, U);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t gi = i0 + k * n4;
      r[k] = gi < n ? Rv[gi * nrhs + c] : 0.0;
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
      D[(i0 + p * n4) * nrhs + c] =
          static_cast<float>(U[0][p] * r[0] + U[1][p] * r[1] + U[2][p] * r[2] + U[3][p] * r[3]);
  }
}

// X[:n] += (V D)[:n] in FP64; R = B
__global__ void rbt_update_kernel(int64_t n, int64_t np, int64_t nrhs, const float* __restrict__ D,
                                  const float* __restrict__ RV, double* __restrict__ X,
                                  int64_t ldx, const double* __restrict__ B, int64_t ldb,
                                  double* __restrict__ Rv, const int* __restrict__ skip) {
  if (skipped(skip)) return;
  const int64_t n4 = np / 4;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n4 * nrhs;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i0 = t / nrhs, c = t % nrhs;
    double V[4][4], d[4];
    rbt_group(RV, np, i0, V);
#pragma unroll
    for (int q = 0; q < 4; ++q) d[q] = static_cast<double>(D[(i0 + q * n4) * nrhs + c]);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int64_t gi = i0 + p * n4;
      if (gi < n) {
        X[gi * ldx + c] += V[p][0] * d[0] + V[p][1] * d[1] + V[p][2] * d[2] + V[p][3] * d[3];
        Rv[gi * nrhs + c] = B[gi * ldb + c];
      }
    }
  }
}

struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
  Timer() {
    hipEventCreate(&a);
    hipEventCreate(&b);
  }
  ~Timer() {
    hipEventDestroy(a);
    hipEventDestroy(b);
  }
  void start(hipStream_t s) { hipEventRecord(a, s); }
  double stop_ms(hipStream_t s) {
    hipEventRecord(b, s);
    hipEventSynchronize(b);
    float ms = 0.f;
    hipEventElapsedTime(&ms, a, b);
    return ms;
  }
};

}  // namespace

// Splits rows x k of src (kcontig: x(r, c) = src[r ld + c], else src[c ld + r]) into
// split_kernel's layout at out (terms planes of rows x Kp) and one scale per 128 rows.
int launch_split(bool bf16, int terms, bool kcontig, int64_t rows, int64_t k, const float* src,
                 int64_t ld, void* out, float* scales, hipStream_t s) {
  if (rows <= 0) return 0;
  if (k <= 0 || k > NBLK || rows > INT32_MAX) {
    set_error("split: %lld x %lld outside k <= %d", (long long)rows, (long long)k, NBLK);
    return DDLA_ERR_INVALID_ARG;
  }
  const int K = (int)k, Kp = (K + 31) / 32 * 32;
  const dim3 grid((unsigned)ceil_div(rows, HB)), block(256);
#define SPL(BF, TM)                                                                              \
  do {                                                                                           \
    typedef typename LowP<BF>::e E;                                                              \
    if (kcontig)                                                                                 \
      hipLaunchKernelGGL((split_kernel<BF, TM, true>), grid, block, 0, s, (int)rows, K, Kp, src, \
                         ld, static_cast<E*>(out), scales);                                      \
    else                                                                                         \
      hipLaunchKernelGGL((split_kernel<BF, TM, false>), grid, block, 0, s, (int)rows, K, Kp,     \
                         src, ld, static_cast<E*>(out), scales);                                 \
  } while (0)
  if (bf16) {
    if (terms == 1) SPL(true, 1); else if (terms == 2) SPL(true, 2); else SPL(true, 3);
  } else {
    if (terms == 1) SPL(false, 1); else if (terms == 2) SPL(false, 2); else SPL(false, 3);
  }
#undef SPL
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

// C (m x n) += alpha A B from split operands: row i of A is a.parts row i (scale block
// i / 128), column j of B is b.parts row j; k <= 128 (padded to Kp in the parts).
int launch_hgemm_parts(bool bf16, int terms, int tri, int64_t m, int64_t n, int64_t k,
                       float alpha, SplitOp a, SplitOp b, float* C, int64_t ldc, hipStream_t s) {
  if (m <= 0 || n <= 0 || k <= 0) return 0;
  const int tm = ceil_div(m, 128), tn = ceil_div(n, 128);
  const int Kp = ((int)k + 31) / 32 * 32;
  const double mn = (tri ? 0.5 : 1.0) * (double)m * n;
  const int64_t tiles = tri ? lower_tile_count(tm, tn) : (int64_t)tm * tn;
  ProfScope prof(bf16 ? "hgemm_mfma_bf16" : "hgemm_mfma_f16",
                 2.0 * mn * k * std::max(1, terms * (terms + 1) / 2),
                 4.0 * 2.0 * mn + 2.0 * terms * ((double)m + n) * k, s);
#define HG(BF, TM, T)                                                                             \
  do {                                                                                            \
    typedef typename LowP<BF>::e E;                                                               \
    hipLaunchKernelGGL((hgemm_kernel<BF, TM, T>), dim3((unsigned)tiles), dim3(256),               \
                       0, s, (int)m, (int)n, Kp, alpha, static_cast<const E*>(a.parts),           \
                       static_cast<const E*>(b.parts), C, ldc, a.scales, b.scales, tm, tn,        \
                       a.plane, b.plane);                                                         \
  } while (0)
#define HG_T(BF, TM)           \
  do {                         \
    if (tri) HG(BF, TM, 1);    \
    else HG(BF, TM, 0);        \
  } while (0)
  if (bf16) {
    if (terms == 1) HG_T(true, 1); else if (terms == 2) HG_T(true, 2); else HG_T(true, 3);
  } else {
    if (terms == 1) HG_T(false, 1); else if (terms == 2) HG_T(false, 2); else HG_T(false, 3);
  }
#undef HG_T
#undef HG
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

// C += alpha * A * B with FP16/BF16 operands and FP32 accumulation (see hgemm_kernel): A m x k
// (k contiguous); B k x n with n contiguous, or (b_kcontig) B(k, j) at B[j ldb + k]; k <= 128.
// Both operands are split into the workspace slot first.
int launch_hgemm(bool bf16, int terms, bool b_kcontig, int tri, int64_t m, int64_t n, int64_t k,
                 float alpha, const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                 int64_t ldc, float* scale_ws, hipStream_t s, int ws_slot) {
  if (m <= 0 || n <= 0 || k <= 0) return 0;
  if (k > NBLK) {
    set_error("hgemm: k = %lld exceeds %d", (long long)k, NBLK);
    return DDLA_ERR_INVALID_ARG;
  }
  terms = std::min(3, std::max(1, terms));
  const int tm = ceil_div(m, 128);
  const int64_t Kp = (k + 31) / 32 * 32;
  void* ws = nullptr;
  DDLA_RETURN_IF(workspace(ws_slot, (size_t)terms * (m + n) * Kp * 2, &ws));
  char* ap = static_cast<char*>(ws);
  char* bp = ap + (size_t)terms * m * Kp * 2;
  DDLA_RETURN_IF(launch_split(bf16, terms, true, m, k, A, lda, ap, scale_ws, s));
  DDLA_RETURN_IF(launch_split(bf16, terms, b_kcontig, n, k, B, ldb, bp, scale_ws + tm, s));
  return launch_hgemm_parts(bf16, terms, tri, m, n, k, alpha, SplitOp{ap, m * Kp, scale_ws},
                            SplitOp{bp, n * Kp, scale_ws + tm}, C, ldc, s);
}

namespace {

// Workspace (WS_MIXED): FP32 working matrix n x n
// Workspace (WS_MIXED2): scales, perm, pivots, D (n x nrhs float), R (n x nrhs double),
//                        row sums, norm words
struct MixedWs {
  int terms;  // low-precision parts per operand in the trailing updates
  float* W;
  float* scales;
  int* perm;
  int32_t* piv;
  float* D;
  double* R;
  double* rowsum;
  double* normA;
  unsigned long long* norms;
  IrState* state;
  float* dinv1;  // inverted 64 x 64 diagonal blocks of the first / second triangular factor
  float* dinv2;
  float* rbt;    // butterfly diagonals: U then V (2 np each)
  int64_t np;    // order of the factored matrix (n, or n rounded up to 4 for butterflies)
};

int mixed_ws(int64_t n, int64_t np, int64_t nrhs, MixedWs* w) {
  void* p1 = nullptr;
  DDLA_RETURN_IF(workspace(WS_MIXED, (size_t)np * np * sizeof(float), &p1));
  w->W = static_cast<float*>(p1);
  w->np = np;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const int64_t nb = (np + 127) / 128 + 1;
  // update scales: main stream, auxiliary stream, two split panels in flight
  const size_t s_scales = al(6 * nb * sizeof(float));
  const size_t s_perm = al(np * sizeof(int));
  const size_t s_piv = al(np * sizeof(int32_t));
  const size_t s_D = al((size_t)np * nrhs * sizeof(float));
  const size_t s_R = al((size_t)n * nrhs * sizeof(double));
  const size_t s_rs = al(n * sizeof(double));
  const size_t s_na = al(sizeof(double));
  const size_t s_no = al(2 * nrhs * sizeof(unsigned long long));
  const size_t s_st = al(sizeof(IrState));
  const size_t s_di = al(trsv_dinv_elems(np) * sizeof(float));
  const size_t s_rb = al(4 * np * sizeof(float));
  void* p2 = nullptr;
  DDLA_RETURN_IF(workspace(
      WS_MIXED2,
      s_scales + s_perm + s_piv + s_D + s_R + s_rs + s_na + s_no + s_st + 2 * s_di + s_rb, &p2));
  char* c = static_cast<char*>(p2);
  w->scales = reinterpret_cast<float*>(c); c += s_scales;
  w->perm = reinterpret_cast<int*>(c); c += s_perm;
  w->piv = reinterpret_cast<int32_t*>(c); c += s_piv;
  w->D = reinterpret_cast<float*>(c); c += s_D;
  w->R = reinterpret_cast<double*>(c); c += s_R;
  w->rowsum = reinterpret_cast<double*>(c); c += s_rs;
  w->normA = reinterpret_cast<double*>(c); c += s_na;
  w->norms = reinterpret_cast<unsigned long long*>(c); c += s_no;
  w->state = reinterpret_cast<IrState*>(c); c += s_st;
  w->dinv1 = reinterpret_cast<float*>(c); c += s_di;
  w->dinv2 = reinterpret_cast<float*>(c); c += s_di;
  w->rbt = reinterpret_cast<float*>(c);
  return 0;
}

// FP32 LU with partial pivoting; trailing updates in prec (FP16/BF16 matrix cores, or
// FP32 matrix cores for DDLA_PREC_FP32).  Fills w.perm.  Returns info via *info_host.
// One trailing update of the mixed LU: U12 = L11^-1 A12 for columns [c0, c0 + nc) of the row
// block, then A22 -= L21 U12 on those columns (FP32, or FP16 / BF16 matrix cores).
// l21: the panel's L21 already split (FP16 / BF16), or null to split it here.
// u12_ready: U12 was already formed (launch_laswp_trsm_slab).
int lu_update_cols(int prec, int terms, int64_t n, int64_t j, int b, int64_t c0, int64_t nc,
                   float* a, int64_t lda, float* scales, hipStream_t s, int split_slot,
                   const SplitOp* l21, bool u12_ready = false) {
  if (nc <= 0) return 0;
  float* panel = a + j * lda + j;
  const int64_t rest = n - j - b;
  float* u12 = a + j * lda + c0;
  if (!u12_ready)
    DDLA_RETURN_IF(launch_trsm_slab<float>(true, true, b, nc, panel, lda, 1, u12, lda, s));
  if (rest <= 0) return 0;
  float* c = a + (j + b) * lda + c0;
  if (l21 && prec != DDLA_PREC_FP32) {
    const bool bf16 = prec == DDLA_PREC_BF16;
    const int64_t Kp = (b + 31) / 32 * 32;
    void* ws = nullptr;
    DDLA_RETURN_IF(workspace(split_slot, (size_t)terms * nc * Kp * 2, &ws));
    DDLA_RETURN_IF(launch_split(bf16, terms, false, nc, b, u12, lda, ws, scales, s));
    return launch_hgemm_parts(bf16, terms, 0, rest, nc, b, -1.f, *l21,
                              SplitOp{ws, nc * Kp, scales}, c, lda, s);
  }
  if (prec == DDLA_PREC_FP32)
    return launch_sgemm(rest, nc, b, -1.f, panel + b * lda, lda, true, u12, lda, true, 1.f, c, lda,
                        0, s);
  return launch_hgemm(prec == DDLA_PREC_BF16, terms, false, 0, rest, nc, b, -1.f, panel + b * lda,
                      lda, u12, lda, c, lda, scales, s, split_slot);
}

// Right-looking blocked LU with partial pivoting on the FP32 working copy (same block
// structure as PLUDecBlock, LU.cs:289-342), or with tournament pivoting in the 128-wide
// panels (tslu.hip; a narrower last panel pivots column by column), with a look-ahead of one
// panel: after panel k, the
// main stream interchanges and updates only the next panel's columns and factors that panel,
// while the auxiliary stream interchanges and updates the remaining columns.  The panel is
// latency-bound (one exchange between workgroups per column), so the bulk of the trailing
// update runs in its shadow.  Ordering, with E0(k) = "panel k done" and E2(k) = "update of
// the remaining columns after panel k done":
//   main: panel(k); wait E2(k-1); swaps + update of the next block; panel(k+1) ...
//   aux:  wait E0(k); swaps on [0, j) and the columns after the next block, their update;
//         record E2(k)
int sgetrf_mixed(int64_t n, int prec, bool tournament, float* a, int64_t lda, MixedWs& w,
                 int32_t* info_host) {
  hipStream_t s = stream();
  hipStream_t s2;
  hipEvent_t* ev;
  DDLA_RETURN_IF(aux_stream(&s2, &ev));
  int32_t* dinfo;
  int *counts, *lists;
  DDLA_RETURN_IF(lu_bookkeeping(n, &dinfo, &counts, &lists));
  DDLA_HIP_CHECK(hipMemsetAsync(dinfo, 0, sizeof(int32_t), s));
  const int64_t nb = (n + NBLK - 1) / NBLK + 1;
  float* scales_aux = w.scales + 2 * nb;
  // FP16 / BF16 updates: the far columns' update splits the panel's L21 once on the auxiliary
  // stream (two copies alternate by panel parity).  The next panel's 128 columns, on the
  // critical path between two panels, take their update in FP32 on the matrix cores instead:
  // no split kernels there, and about 1/64 of the update flops at n = 16384.
  const bool share = prec != DDLA_PREC_FP32;
  const size_t pslot = (size_t)w.terms * n * NBLK * 2;
  char* pbuf = nullptr;
  if (share) {
    void* t = nullptr;
    DDLA_RETURN_IF(workspace(WS_SPLITP, 2 * pslot, &t));
    pbuf = static_cast<char*>(t);
  }
  // the auxiliary stream starts after everything queued so far on the main stream
  DDLA_HIP_CHECK(hipEventRecord(ev[0], s));
  DDLA_HIP_CHECK(hipStreamWaitEvent(s2, ev[0], 0));
  bool pending = false;  // an update on the auxiliary stream that the main stream must await
  int k = 0;
  for (int64_t j = 0; j < n; j += NBLK, ++k) {
    const int b = (int)std::min<int64_t>(NBLK, n - j);
    float* panel = a + j * lda + j;
    int* list_k = lists + (size_t)k * 4 * NBLK;
    if (tournament && tslu_supported(n - j, b))
      DDLA_RETURN_IF(launch_tslu_panel(n - j, b, panel, lda, dinfo, (int)j, list_k, counts + k, s));
    else
      DDLA_RETURN_IF(launch_getf2_panel_ex<float>(n - j, b, panel, lda, w.piv + j, dinfo, (int)j,
                                                  list_k, counts + k, s));
    const int64_t nxt = std::min<int64_t>(NBLK, n - j - b);  // the next panel's width
    const int64_t far0 = j + b + nxt, far = n - far0;        // columns after the next panel
    if (far > 0) {
      DDLA_HIP_CHECK(hipEventRecord(ev[0], s));  // E0(k)
    }
    if (pending) {
      DDLA_HIP_CHECK(hipStreamWaitEvent(s, ev[1], 0));  // E2(k-1): next block fully updated
      pending = false;
    }
    // main stream: swaps on the next block (and on [0, j) when no far columns are left), its
    // update (FP32)
    // (with far columns left, the next block's swaps and its U12 are one launch)
    if (far > 0)
      DDLA_RETURN_IF(launch_laswp_trsm_slab<float>(b, nxt, panel, lda, a + j * lda + j + b, lda,
                                                   list_k, counts + k, s));
    else
      DDLA_RETURN_IF(launch_laswp_list<float>(a + j * lda, lda, 0, j + b + nxt, j, j + b, list_k,
                                              counts + k, 2 * NBLK, s));
    DDLA_RETURN_IF(lu_update_cols(DDLA_PREC_FP32, w.terms, n, j, b, j + b, nxt, a, lda,
                                  w.scales, s, WS_SPLIT, nullptr, far > 0));
    // auxiliary stream: swaps on [0, j) and the far columns (off the critical path: nothing
    // reads the finished columns until the end), update of the far columns
    if (far > 0) {
      DDLA_HIP_CHECK(hipStreamWaitEvent(s2, ev[0], 0));
      DDLA_RETURN_IF(launch_laswp_list<float>(a + j * lda, lda, 0, n, j, far0, list_k, counts + k,
                                              2 * NBLK, s2));
      SplitOp l21{};
      if (share) {
        const int64_t rest = n - j - b, Kp = (b + 31) / 32 * 32;
        float* sc = w.scales + (4 + (k & 1)) * nb;
        char* out = pbuf + (k & 1) * pslot;
        DDLA_RETURN_IF(launch_split(prec == DDLA_PREC_BF16, w.terms, true, rest, b,
                                    panel + b * lda, lda, out, sc, s2));
        l21 = SplitOp{out, rest * Kp, sc};
      }
      DDLA_RETURN_IF(lu_update_cols(prec, w.terms, n, j, b, far0, far, a, lda, scales_aux, s2,
                                    WS_SPLIT2, share ? &l21 : nullptr));
      DDLA_HIP_CHECK(hipEventRecord(ev[1], s2));  // E2(k)
      pending = true;
    }
  }
  if (pending) DDLA_HIP_CHECK(hipStreamWaitEvent(s, ev[1], 0));
  DDLA_RETURN_IF(launch_perm_compose((int)n, k, NBLK, lists, counts, w.perm, s));
  DDLA_HIP_CHECK(hipMemcpyAsync(info_host, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DDLA_HIP_CHECK(hipStreamSynchronize(s));
  return check_abort(panel_abort_word(nullptr), s, "mixed LU panel");
}

// FP32 LU without pivoting of the butterfly-transformed matrix (same block structure as
// LUDecBlock, LU.cs:169-186); trailing updates in prec.
int sgetrf_nopiv_mixed(int64_t n, int prec, MixedWs& w, int32_t* info_host) {
  hipStream_t s = stream();
  int32_t* dinfo;
  int *counts, *lists;
  DDLA_RETURN_IF(lu_bookkeeping(1, &dinfo, &counts, &lists));
  DDLA_HIP_CHECK(hipMemsetAsync(dinfo, 0, sizeof(int32_t), s));
  float* a = w.W;
  const int64_t lda = n;
  for (int64_t j = 0; j < n; j += NBLK) {
    const int b = (int)std::min<int64_t>(NBLK, n - j);
    float* a11 = a + j * lda + j;
    DDLA_RETURN_IF(launch_getf2_nopiv<float>(b, a11, lda, dinfo, (int)j, s));
    const int64_t rest = n - j - b;
    if (rest > 0) {
      DDLA_RETURN_IF(launch_trsm_slab<float>(true, true, b, rest, a11, lda, 1, a11 + b, lda, s));
      DDLA_RETURN_IF(launch_trsm_rlt_panel<float>(b, rest, a11, lda, a11 + b * lda, lda, s, true));
      if (prec == DDLA_PREC_FP32) {
        DDLA_RETURN_IF(launch_sgemm(rest, rest, b, -1.f, a11 + b * lda, lda, true, a11 + b, lda, true,
                                    1.f, a11 + b * lda + b, lda, 0, s));
      } else {
        DDLA_RETURN_IF(launch_hgemm(prec == DDLA_PREC_BF16, w.terms, false, 0, rest, rest, b, -1.f,
                                    a11 + b * lda, lda, a11 + b, lda, a11 + b * lda + b, lda,
                                    w.scales, s));
      }
    }
  }
  DDLA_HIP_CHECK(hipMemcpyAsync(info_host, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DDLA_HIP_CHECK(hipStreamSynchronize(s));
  return 0;
}

// FP32 lower Cholesky; trailing SYRK in prec, with the look-ahead of dpotrf_internal: the
// next block column is updated on the main stream, the rest of the trailing triangle on the
// auxiliary stream (its own split workspace and scales), in the shadow of the next diagonal
// block and L21.
int spotrf_mixed(int64_t n, int prec, MixedWs& w, int32_t* info_host) {
  hipStream_t s = stream();
  hipStream_t s2;
  hipEvent_t* ev;
  DDLA_RETURN_IF(aux_stream(&s2, &ev));
  int32_t* dinfo;
  int *counts, *lists;
  DDLA_RETURN_IF(lu_bookkeeping(1, &dinfo, &counts, &lists));
  DDLA_HIP_CHECK(hipMemsetAsync(dinfo, 0, sizeof(int32_t), s));
  float* a = w.W;
  const int64_t lda = n;
  const int64_t nb = (n + NBLK - 1) / NBLK + 1;
  float* scales_aux = w.scales + 2 * nb;
  auto syrk = [&](int64_t m, int64_t nc, int b, const float* A, float* C, float* sc,
                  hipStream_t st, int slot) -> int {
    if (prec == DDLA_PREC_FP32)
      return launch_sgemm(m, nc, b, -1.f, A, lda, true, A, lda, false, 1.f, C, lda, 1, st);
    return launch_hgemm(prec == DDLA_PREC_BF16, w.terms, true, 1, m, nc, b, -1.f, A, lda, A, lda,
                        C, lda, sc, st, slot);
  };
  // FP16 / BF16: L21 is split once per block column and serves as both operands of both
  // streams' updates (two copies alternate, as in sgetrf_mixed)
  const bool share = prec != DDLA_PREC_FP32;
  const size_t pslot = (size_t)w.terms * n * NBLK * 2;
  char* pbuf = nullptr;
  if (share) {
    void* t = nullptr;
    DDLA_RETURN_IF(workspace(WS_SPLITP, 2 * pslot, &t));
    pbuf = static_cast<char*>(t);
  }
  DDLA_HIP_CHECK(hipEventRecord(ev[0], s));
  DDLA_HIP_CHECK(hipStreamWaitEvent(s2, ev[0], 0));
  bool pending = false;
  int k = 0;
  for (int64_t j = 0; j < n; j += NBLK, ++k) {
    const int b = (int)std::min<int64_t>(NBLK, n - j);
    float* a11 = a + j * lda + j;
    DDLA_RETURN_IF(launch_potf2_ex<float>(b, a11, lda, dinfo, (int)j, s));
    const int64_t rest = n - j - b;
    if (rest <= 0) break;
    float* a21 = a11 + b * lda;
    DDLA_RETURN_IF(launch_trsm_rlt_panel<float>(b, rest, a11, lda, a21, lda, s, false));
    const int64_t nxt = std::min<int64_t>(NBLK, rest), far = rest - nxt;
    SplitOp l21{};
    const int64_t Kp = (b + 31) / 32 * 32;
    if (share) {
      float* sc = w.scales + (4 + (k & 1)) * nb;
      char* out = pbuf + (k & 1) * pslot;
      DDLA_RETURN_IF(launch_split(prec == DDLA_PREC_BF16, w.terms, true, rest, b, a21, lda, out,
                                  sc, s));
      l21 = SplitOp{out, rest * Kp, sc};
    }
    if (far > 0) DDLA_HIP_CHECK(hipEventRecord(ev[0], s));  // E0(k)
    if (pending) {
      DDLA_HIP_CHECK(hipStreamWaitEvent(s, ev[1], 0));  // E2(k-1)
      pending = false;
    }
    if (share)
      DDLA_RETURN_IF(launch_hgemm_parts(prec == DDLA_PREC_BF16, w.terms, 1, rest, nxt, b, -1.f,
                                        l21, l21, a21 + b, lda, s));
    else
      DDLA_RETURN_IF(syrk(rest, nxt, b, a21, a21 + b, w.scales, s, WS_SPLIT));
    if (far > 0) {
      DDLA_HIP_CHECK(hipStreamWaitEvent(s2, ev[0], 0));
      float* af = a21 + nxt * lda;
      if (share) {
        // rows nxt.. of the split L21: nxt = 128 here, one whole scale block
        const SplitOp lf{static_cast<const char*>(l21.parts) + (size_t)nxt * Kp * 2, l21.plane,
                         l21.scales + nxt / HB};
        DDLA_RETURN_IF(launch_hgemm_parts(prec == DDLA_PREC_BF16, w.terms, 1, far, far, b, -1.f,
                                          lf, lf, af + b + nxt, lda, s2));
      } else {
        DDLA_RETURN_IF(syrk(far, far, b, af, af + b + nxt, scales_aux, s2, WS_SPLIT2));
      }
      DDLA_HIP_CHECK(hipEventRecord(ev[1], s2));  // E2(k)
      pending = true;
    }
  }
  if (pending) DDLA_HIP_CHECK(hipStreamWaitEvent(s, ev[1], 0));
  DDLA_HIP_CHECK(hipMemcpyAsync(info_host, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DDLA_HIP_CHECK(hipStreamSynchronize(s));
  return 0;
}

// ||A||_inf into w.normA (device; read there by the convergence test)
int norm_inf(int64_t n, const double* A, int64_t lda, MixedWs& w) {
  hipStream_t s = stream();
  hipLaunchKernelGGL(row_abs_sum_kernel, dim3(ceil_div(n, 4)), dim3(256), 0, s, n, A, lda, w.rowsum);
  hipLaunchKernelGGL(max_kernel, dim3(1), dim3(1024), 0, s, n, w.rowsum, w.normA);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

// Classic refinement. lu: solve with P, L (unit) and U; else L L^T.
enum FactorKind { LU_PIVOT = 0, LU_BUTTERFLY = 1, CHOLESKY = 2, LU_TOURNAMENT = 3 };

// Classic refinement: X += M^-1 R, R = B - A X, with M the FP32 factor of kind `kind`.
// Stops at backward error <= tol, after max_iters steps, or when three steps in a row fail
// to lower the backward error by 10% (stagnation).  The test runs on the GPU after every step
// (ir_norms_check_kernel); the host queues STEPS steps at a time, one group ahead of the state
// it reads (steps queued after convergence do nothing).  With more than 8 right-hand sides the
// triangular solves and the residual are library GEMMs without the skip word, so the host
// reads the state after every step there before it queues the next.
int refine(FactorKind kind, int64_t n, int64_t nrhs, const double* A, int64_t lda, const double* B,
           int64_t ldb, double* X, int64_t ldx, MixedWs& w, double tol, int max_iters, int* iters,
           double* berr_out) {
  hipStream_t s = stream();
  const int64_t tot = n * nrhs;
  const int64_t np = w.np;
  const int64_t totp = (np / 4) * nrhs;
  const float* RU = w.rbt;
  const float* RV = w.rbt + 2 * np;
  const bool few = nrhs <= 8;
  const int* skip = few ? &w.state->done : nullptr;
  // Steps are queued in groups.  With few right-hand sides every kernel of a step is a no-op
  // once the device-side test has converged, so the next group is queued before the host reads
  // the state of the current one (the GPU does not wait for the host between groups; at most
  // one group of no-op steps follows convergence).  State and the solves' abort word reach
  // the host by asynchronous copies into pinned memory, one slot per group parity.
  const int STEPS = few ? 2 : 1;
  // X = 0, R = B
  DDLA_RETURN_IF(launch_dscale2d(n, nrhs, 0.0, X, ldx, 1, 0, s));
  DDLA_RETURN_IF(launch_dcopy2d(n, nrhs, B, ldb, 1, w.R, nrhs, s));
  hipLaunchKernelGGL(ir_state_init_kernel, dim3(1), dim3(1), 0, s, w.state, w.norms,
                     (int)(2 * nrhs));
  DDLA_HIP_CHECK(hipGetLastError());
  struct Slot {
    IrState st;
    unsigned abort;
  };
  static_assert(2 * sizeof(Slot) <= PINNED_BYTES, "pinned slots");
  void* pin = nullptr;
  DDLA_RETURN_IF(pinned_host(&pin));
  Slot* slots = static_cast<Slot*>(pin);
  hipStream_t s2;
  hipEvent_t* ev;
  DDLA_RETURN_IF(aux_stream(&s2, &ev));  // ev[2], ev[3]: the group copies' completion
  int queued = 0;
  auto queue_group = [&](int gi) -> int {
    const int group = std::min(STEPS, max_iters - queued);
    for (int q = 0; q < group; ++q) {
      // D = M^-1 R in FP32
      if (kind == LU_BUTTERFLY)
        hipLaunchKernelGGL(rbt_vec_kernel, dim3(grid_for(totp)), dim3(256), 0, s, n, np, nrhs,
                           w.R, RU, w.D, skip);
      else
        hipLaunchKernelGGL(gather_d2s_kernel, dim3(grid_for(tot)), dim3(256), 0, s, n, nrhs, w.R,
                           nrhs, kind == LU_PIVOT || kind == LU_TOURNAMENT ? w.perm : nullptr,
                           w.D, skip);
      DDLA_HIP_CHECK(hipGetLastError());
      const bool lu = kind != CHOLESKY;
      // triangular solves with the inverted diagonal blocks (trsv.hip): the persistent
      // latency-bound solve for up to 8 right-hand sides, GEMMs on the same blocks beyond
      if (few) {
        DDLA_RETURN_IF(trsv_solve<float>(true, false, np, (int)nrhs, w.W, np, w.dinv1, w.D, nrhs,
                                         s, skip));
        DDLA_RETURN_IF(trsv_solve<float>(false, !lu, np, (int)nrhs, w.W, np, w.dinv2, w.D, nrhs,
                                         s, skip));
      } else {
        DDLA_RETURN_IF(trsv_solve_blocked(true, false, np, (int)nrhs, w.W, np, w.dinv1, w.D, nrhs, s));
        DDLA_RETURN_IF(trsv_solve_blocked(false, !lu, np, (int)nrhs, w.W, np, w.dinv2, w.D, nrhs, s));
      }
      // X += D (FP64, through V for the butterfly form); R = B - A X (FP64)
      if (kind == LU_BUTTERFLY)
        hipLaunchKernelGGL(rbt_update_kernel, dim3(grid_for(totp)), dim3(256), 0, s, n, np, nrhs,
                           w.D, RV, X, ldx, B, ldb, w.R, skip);
      else
        hipLaunchKernelGGL(update_x_kernel, dim3(grid_for(tot)), dim3(256), 0, s, n, nrhs, w.D, X,
                           ldx, B, ldb, w.R, skip);
      DDLA_HIP_CHECK(hipGetLastError());
      if (few) {
        ProfScope prof("dgemv_resid_f64", 2.0 * n * n * nrhs, 8.0 * n * n, s);
        const dim3 g((unsigned)ceil_div(n, 16));  // 4 waves x 4 rows per workgroup
        if (nrhs == 1)
          hipLaunchKernelGGL(dgemv_resid_kernel<1>, g, dim3(256), 0, s, n, A, lda, X, ldx, w.R,
                             nrhs, (int)nrhs, skip);
        else
          hipLaunchKernelGGL(dgemv_resid_kernel<8>, g, dim3(256), 0, s, n, A, lda, X, ldx, w.R,
                             nrhs, (int)nrhs, skip);
      } else {
        DDLA_RETURN_IF(launch_dgemm(n, nrhs, n, -1.0, A, lda, true, X, ldx, true, 1.0, w.R, nrhs,
                                    0, s));
      }
      hipLaunchKernelGGL(ir_norms_check_kernel,
                         dim3((unsigned)std::min<int64_t>(ceil_div(n, 256), 64), (unsigned)nrhs),
                         dim3(256), 0, s, n, nrhs, w.R, X, ldx, w.norms, w.normA, tol, max_iters,
                         w.state);
      DDLA_HIP_CHECK(hipGetLastError());
    }
    queued += group;
    Slot* sl = slots + (gi & 1);
    DDLA_HIP_CHECK(hipMemcpyAsync(&sl->st, w.state, sizeof(IrState), hipMemcpyDeviceToHost, s));
    if (few)
      DDLA_HIP_CHECK(hipMemcpyAsync(&sl->abort, trsv_abort_word(), sizeof(unsigned),
                                    hipMemcpyDeviceToHost, s));
    DDLA_HIP_CHECK(hipEventRecord(ev[2 + (gi & 1)], s));
    return 0;
  };
  IrState st{};
  DDLA_RETURN_IF(queue_group(0));
  int last = 0;  // the last group queued
  for (int g = 0;; ++g) {
    if (few && last == g && queued < max_iters) DDLA_RETURN_IF(queue_group(last = g + 1));
    DDLA_HIP_CHECK(hipEventSynchronize(ev[2 + (g & 1)]));
    st = slots[g & 1].st;
    if (few && slots[g & 1].abort)
      DDLA_RETURN_IF(check_abort(trsv_abort_word(), s, "refinement triangular solve"));
    if (st.done) break;
    if (last == g) {  // nothing queued beyond this group
      if (queued >= max_iters) break;
      DDLA_RETURN_IF(queue_group(last = g + 1));
    }
  }
  *iters = st.iters;
  *berr_out = st.berr;
  return 0;
}

// max over columns of ||x_j - xref_j||_inf / ||xref_j||_inf
int forward_error(int64_t n, int64_t nrhs, const double* X, int64_t ldx, const double* Xr,
                  int64_t ldr, MixedWs& w, double* out) {
  hipStream_t s = stream();
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(256), 0, s, 2 * nrhs, w.norms, nullptr);
  hipLaunchKernelGGL(fwd_err_norms_kernel,
                     dim3((unsigned)std::min<int64_t>(ceil_div(n, 256), 64), (unsigned)nrhs),
                     dim3(256), 0, s, n, X, ldx, Xr, ldr, w.norms);
  DDLA_HIP_CHECK(hipGetLastError());
  std::vector<unsigned long long> h(2 * nrhs);
  DDLA_HIP_CHECK(hipMemcpyAsync(h.data(), w.norms, h.size() * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, s));
  DDLA_HIP_CHECK(hipStreamSynchronize(s));
  double worst = 0.0;
  for (int64_t j = 0; j < nrhs; ++j) {
    const double d = __builtin_bit_cast(double, h[2 * j]);
    const double r = __builtin_bit_cast(double, h[2 * j + 1]);
    const double e = r > 0.0 ? d / r : (d == 0.0 ? 0.0 : INFINITY);
    worst = std::max(worst, std::isnan(e) ? INFINITY : e);
  }
  *out = worst;
  return 0;
}

int mpir(bool lu, int64_t n, int64_t nrhs, const double* a, int64_t lda, const double* b,
         int64_t ldb, double* x, int64_t ldx, const ddla_mpir_opts* opts, ddla_mpir_info* info) {
  hipStream_t s = stream();
  ddla_mpir_info out{};
  int prec = opts ? opts->factor_prec : DDLA_PREC_FP16;
  const int max_iters = (opts && opts->max_iters > 0) ? opts->max_iters : 50;
  const double tol = (opts && opts->tol > 0) ? opts->tol : 1e-14;
  if (prec != DDLA_PREC_FP16 && prec != DDLA_PREC_BF16 && prec != DDLA_PREC_FP32) {
    set_error("factor precision %d not supported by the mixed solver", prec);
    return DDLA_ERR_UNSUPPORTED;
  }
  if (opts && opts->gmres != 0) {
    set_error("GMRES refinement is not available (opts->gmres must be 0)");
    return DDLA_ERR_UNSUPPORTED;
  }
  FactorKind kind = lu ? LU_PIVOT : CHOLESKY;
  if (lu && opts && opts->pivoting == DDLA_PIVOT_BUTTERFLY) kind = LU_BUTTERFLY;
  if (lu && opts && opts->pivoting == DDLA_PIVOT_TOURNAMENT) kind = LU_TOURNAMENT;
  const int64_t np = (kind == LU_BUTTERFLY) ? (n + 3) / 4 * 4 : n;
  MixedWs w;
  DDLA_RETURN_IF(mixed_ws(n, np, nrhs, &w));
  // default split: FP16 x2 (22 bits) and BF16 x3 (24 bits) reach FP32 product accuracy
  w.terms = (opts && opts->terms > 0) ? std::min(opts->terms, 3)
                                      : (prec == DDLA_PREC_BF16 ? 3 : 2);
  // ||A||_inf: in the same pass as the FP32 copy of A, except for the butterfly LU (whose
  // factor does not start from that copy)
  const bool fuse_norm = kind != LU_BUTTERFLY;
  if (!fuse_norm) DDLA_RETURN_IF(norm_inf(n, a, lda, w));
  Timer t;
  double factor_ms = 0.0, refine_ms = 0.0;
  int iters = 0;
  double berr = INFINITY;
  int32_t finfo = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    t.start(s);
    finfo = 0;
    if (kind == LU_BUTTERFLY) {
      // random diagonals (fixed seed: results are reproducible), then W = U^T A V
      hipLaunchKernelGGL(rbt_diag_kernel, dim3(grid_for(4 * np)), dim3(256), 0, s, 4 * np,
                         (uint64_t)0xDD1ABF00ull, w.rbt);
      hipLaunchKernelGGL(rbt_matrix_kernel, dim3(grid_for((np / 4) * (np / 4))), dim3(256), 0, s,
                         n, np, a, lda, w.rbt, w.rbt + 2 * np, w.W);
      DDLA_HIP_CHECK(hipGetLastError());
      DDLA_RETURN_IF(sgetrf_nopiv_mixed(np, prec, w, &finfo));
    } else {
      if (attempt == 0 && fuse_norm) {
        hipLaunchKernelGGL(d2s_rowsum_kernel, dim3(ceil_div(n, 4)), dim3(256), 0, s, n, a, lda,
                           w.W, n, w.rowsum);
        hipLaunchKernelGGL(max_kernel, dim3(1), dim3(1024), 0, s, n, w.rowsum, w.normA);
      } else {
        hipLaunchKernelGGL(d2s_kernel, dim3(grid_for(n * n)), dim3(256), 0, s, n, n, a, lda, w.W,
                           n);
      }
      DDLA_HIP_CHECK(hipGetLastError());
      if (kind == LU_PIVOT || kind == LU_TOURNAMENT) {
        DDLA_RETURN_IF(sgetrf_mixed(n, prec, kind == LU_TOURNAMENT, w.W, n, w, &finfo));
      } else {
        DDLA_RETURN_IF(spotrf_mixed(n, prec, w, &finfo));
        if (finfo > 0 && prec != DDLA_PREC_FP32) {
          // the low-precision factor lost positive definiteness: FP32 factor instead
          out.fallback = 1;
          hipLaunchKernelGGL(d2s_kernel, dim3(grid_for(n * n)), dim3(256), 0, s, n, n, a, lda,
                             w.W, n);
          DDLA_RETURN_IF(spotrf_mixed(n, DDLA_PREC_FP32, w, &finfo));
        }
      }
    }
    if (finfo == 0) {
      // inverted diagonal blocks of the two triangular factors for the refinement solves
      if (kind != CHOLESKY) {  // w.np: the butterfly's padded order, or n after the fallback
        DDLA_RETURN_IF(trsv_prepare<float>(true, true, false, w.np, w.W, w.np, w.dinv1, s));
        DDLA_RETURN_IF(trsv_prepare<float>(false, false, false, w.np, w.W, w.np, w.dinv2, s));
      } else {
        DDLA_RETURN_IF(trsv_prepare<float>(true, false, false, n, w.W, n, w.dinv1, s));
        DDLA_RETURN_IF(trsv_prepare<float>(false, false, true, n, w.W, n, w.dinv2, s));
      }
    }
    factor_ms += t.stop_ms(s);
    if (finfo == 0) {
      t.start(s);
      DDLA_RETURN_IF(refine(kind, n, nrhs, a, lda, b, ldb, x, ldx, w, tol, max_iters, &iters,
                            &berr));
      refine_ms += t.stop_ms(s);
    }
    // a butterfly or tournament factor that broke down or does not refine: redo with partial
    // pivoting
    if ((kind == LU_BUTTERFLY || kind == LU_TOURNAMENT) && (finfo != 0 || !(berr <= tol))) {
      out.fallback = kind == LU_BUTTERFLY ? 2 : 3;
      kind = LU_PIVOT;
      w.np = n;
      continue;
    }
    break;
  }
  out.factor_ms = factor_ms;
  out.refine_ms = refine_ms;
  if (finfo > 0) {
    out.status = finfo;
    if (info) *info = out;
    return finfo;
  }
  out.iters = iters;
  out.backward_error = berr;
  out.forward_error = -1.0;
  if (opts && opts->x_ref)
    DDLA_RETURN_IF(forward_error(n, nrhs, x, ldx, opts->x_ref, opts->ld_xref, w, &out.forward_error));
  out.status = (berr <= tol) ? 0 : DDLA_ERR_NOT_CONVERGED;
  if (info) *info = out;
  if (out.status) set_error("refinement stopped at backward error %.3e after %d steps", berr, iters);
  return out.status;
}

// FP32 LU of a device matrix in place with the mixed solver's trailing updates; perm[i] =
// source row of row i.
int sgetrf_mp(int64_t n, float* a, int64_t lda, int prec, int pivoting, int32_t* perm) {
  MixedWs w;
  DDLA_RETURN_IF(mixed_ws(n, n, 1, &w));
  w.terms = prec == DDLA_PREC_BF16 ? 3 : 2;
  int32_t finfo = 0;
  DDLA_RETURN_IF(sgetrf_mixed(n, prec, pivoting == DDLA_PIVOT_TOURNAMENT, a, lda, w, &finfo));
  DDLA_HIP_CHECK(hipMemcpyAsync(perm, w.perm, n * sizeof(int32_t), hipMemcpyDeviceToDevice,
                                stream()));
  DDLA_HIP_CHECK(hipStreamSynchronize(stream()));
  return finfo;
}

}  // namespace

}  // namespace ddla

using namespace ddla;

extern "C" {

int ddla_dgesv_mpir_dev(ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t lda,
                        const double* b, ddla_inc_t ldb, double* x, ddla_inc_t ldx,
                        const ddla_mpir_opts* opts, ddla_mpir_info* info) {
  DDLA_RETURN_IF(ensure_init());
  DDLA_ARG_CHECK(n >= 0 && nrhs >= 0 && lda >= n && ldb >= nrhs && ldx >= nrhs,
                 "bad gesv arguments");
  DDLA_ARG_CHECK(!opts || !opts->x_ref || opts->ld_xref >= nrhs, "ld_xref below nrhs");
  if (n == 0 || nrhs == 0) return 0;
  DDLA_ARG_CHECK(n <= 32768, "n = %lld above the supported 32768", (long long)n);
  return mpir(true, n, nrhs, a, lda, b, ldb, x, ldx, opts, info);
}

int ddla_sgetrf_mp_dev(ddla_dim_t n, float* a, ddla_inc_t lda, int32_t prec, int32_t pivoting,
                       int32_t* perm) {
  DDLA_RETURN_IF(ensure_init());
  DDLA_ARG_CHECK(n >= 0 && lda >= n, "bad sgetrf arguments");
  DDLA_ARG_CHECK(prec == DDLA_PREC_FP32 || prec == DDLA_PREC_FP16 || prec == DDLA_PREC_BF16,
                 "trailing-update precision %d not supported", (int)prec);
  DDLA_ARG_CHECK(pivoting == DDLA_PIVOT_PARTIAL || pivoting == DDLA_PIVOT_TOURNAMENT,
                 "pivoting %d not supported by the FP32 factor", (int)pivoting);
  if (n == 0) return 0;
  DDLA_ARG_CHECK(n <= 32768, "n = %lld above the supported 32768", (long long)n);
  return sgetrf_mp(n, a, lda, prec, pivoting, perm);
}

int ddla_sgetf2_dev(ddla_dim_t m, ddla_dim_t nb, float* p, ddla_inc_t ldp, int32_t* piv) {
  DDLA_RETURN_IF(ensure_init());
  DDLA_ARG_CHECK(nb >= 1 && nb <= NBLK && m >= nb && ldp >= nb && p && piv,
                 "bad sgetf2 arguments (m %lld, nb %lld, ld %lld)", (long long)m, (long long)nb,
                 (long long)ldp);
  DDLA_ARG_CHECK(m <= 32768, "m = %lld above the supported 32768", (long long)m);
  hipStream_t s = stream();
  int32_t* dinfo;
  int *counts, *lists;
  DDLA_RETURN_IF(lu_bookkeeping(m, &dinfo, &counts, &lists));
  DDLA_RETURN_IF(launch_fill_u32(dinfo, 1, 0u, s));
  DDLA_RETURN_IF(
      launch_getf2_panel_ex<float>(m, (int)nb, p, ldp, piv, dinfo, 0, lists, counts, s));
  int32_t h = 0;
  DDLA_HIP_CHECK(hipMemcpyAsync(&h, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DDLA_RETURN_IF(check_abort(panel_abort_word(nullptr), s, "FP32 panel"));
  return h;
}

int ddla_dgetf2_dev(ddla_dim_t m, ddla_dim_t nb, double* p, ddla_inc_t ldp, int32_t* piv) {
  DDLA_RETURN_IF(ensure_init());
  DDLA_ARG_CHECK(nb >= 1 && nb <= NBLK && m >= nb && ldp >= nb && p && piv,
                 "bad dgetf2 arguments (m %lld, nb %lld, ld %lld)", (long long)m, (long long)nb,
                 (long long)ldp);
  DDLA_ARG_CHECK(m <= 32768, "m = %lld above the supported 32768", (long long)m);
  hipStream_t s = stream();
  int32_t* dinfo;
  int *counts, *lists;
  DDLA_RETURN_IF(lu_bookkeeping(m, &dinfo, &counts, &lists));
  DDLA_RETURN_IF(launch_fill_u32(dinfo, 1, 0u, s));
  DDLA_RETURN_IF(
      launch_getf2_panel_ex<double>(m, (int)nb, p, ldp, piv, dinfo, 0, lists, counts, s));
  int32_t h = 0;
  DDLA_HIP_CHECK(hipMemcpyAsync(&h, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DDLA_RETURN_IF(check_abort(panel_abort_word(nullptr), s, "FP64 panel"));
  return h;
}

int ddla_dposv_mpir_dev(ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t lda,
                        const double* b, ddla_inc_t ldb, double* x, ddla_inc_t ldx,
                        const ddla_mpir_opts* opts, ddla_mpir_info* info) {
  DDLA_RETURN_IF(ensure_init());
  DDLA_ARG_CHECK(n >= 0 && nrhs >= 0 && lda >= n && ldb >= nrhs && ldx >= nrhs,
                 "bad posv arguments");
  DDLA_ARG_CHECK(!opts || !opts->x_ref || opts->ld_xref >= nrhs, "ld_xref below nrhs");
  if (n == 0 || nrhs == 0) return 0;
  return mpir(false, n, nrhs, a, lda, b, ldb, x, ldx, opts, info);
}

}  // extern "C"
