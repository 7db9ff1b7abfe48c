// FP64 / FP32 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4_f64,
// v_mfma_f32_16x16x4_f32).
//
// Replaces the managed BLIS-style GEMM (DDLA/BLAS/Managed/3m/GeMM.cs:14-73, packing
// Packing.cs:30-101, macro-kernel MacroKernel.cs:24-97, micro-kernel
// DDLA/Einsum/GEMMKernel.cs:36-240) and the BLIS dgemm it stands in for
// (DDLA/BLAS/Native/3m.cs:12-24).  Design:
//   * one workgroup (256 lanes = 4 wave64) computes a BM x BN tile of row-major C;
//     the 4 waves form a 2 x 2 grid and each wave owns (BM/2) x (BN/2) = TM x TN MFMA
//     16x16 tiles held in accumulator registers;
//   * K advances in steps of 16 (32 for the 64 x 64 tiles): the next A/B slices are fetched
//     from HBM into registers while the matrix cores work on the current slices in LDS (double
//     buffered LDS, one barrier per step) -- this replaces the CPU path's explicit
//     packing pass;
//   * LDS rows are padded by two elements so the MFMA operand reads (16 lanes = 16 rows
//     of A at one k, or 16 columns of B at one k) hit distinct banks;
//   * workgroup ids are remapped so every XCD (own L2) walks a contiguous band of a
//     grouped tile order, keeping the A row panels and B column panels it reuses in its
//     own L2;
//   * epilogue C = alpha*acc + beta*C in registers (beta == 0 never reads C, matching
//     Managed/1m.cs:265-269), optional lower/upper triangle mask for SYRK.
// Operand layout of both 16x16x4 instructions: lane l supplies A[l%16][l/16] and
// B[l/16][l%16].  The result layouts differ (measured on gfx950): the FP32 form returns
// D[4*(l/16)+r][l%16] in register r, the FP64 form D[l/16+4*r][l%16].
#include "ddla_internal.hpp"

namespace ddla {
namespace {

template <typename T>
struct Vec;
template <>
struct Vec<double> {
  typedef double v2 __attribute__((ext_vector_type(2)));
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mfma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row within the 16x16 tile of result register r of lane group lk = lane/16
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <>
struct Vec<float> {
  typedef float v2 __attribute__((ext_vector_type(2)));
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mfma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

template <typename T, int BM_, int BN_, int BK_ = 16>
struct Cfg {
  typedef T E;
  typedef typename Vec<T>::v2 v2;
  typedef typename Vec<T>::v4 v4;
  static constexpr int BM = BM_, BN = BN_, BK = BK_;
  static constexpr int WM = BM / 2, WN = BN / 2;
  static constexpr int TM = WM / 16, TN = WN / 16;
  static constexpr int PA = BK + 2;  // LDS row stride of A tile ([m][k])
  static constexpr int PB = BN + 2;  // LDS row stride of B tile ([k][n])
  static constexpr int ACH = BM * BK / 2 / 256;  // 2-element chunks per lane
  static constexpr int BCH = BK * BN / 2 / 256;
  static constexpr int AS = BM * PA, BS = BK * PB;
};

template <class C, bool AK>
__device__ __forceinline__ void load_a(typename C::v2 (&r)[C::ACH],
                                       const typename C::E* __restrict__ A, int64_t lda, int M,
                                       int K, int m0, int k0, bool vec) {
  typedef typename C::E E;
  typedef typename C::v2 v2;
#pragma unroll
  for (int q = 0; q < C::ACH; ++q) {
    const int v = threadIdx.x + 256 * q;
    int gi, gk;
    if (AK) {
      gi = m0 + v / (C::BK / 2);
      gk = k0 + (v % (C::BK / 2)) * 2;
      const E* p = A + (int64_t)gi * lda + gk;
      if (vec && gi < M && gk + 1 < K) {
        r[q] = *reinterpret_cast<const v2*>(p);
      } else {
        r[q].x = (gi < M && gk < K) ? p[0] : E(0);
        r[q].y = (gi < M && gk + 1 < K) ? p[1] : E(0);
      }
    } else {
      gk = k0 + v / (C::BM / 2);
      gi = m0 + (v % (C::BM / 2)) * 2;
      const E* p = A + gi + (int64_t)gk * lda;
      if (vec && gk < K && gi + 1 < M) {
        r[q] = *reinterpret_cast<const v2*>(p);
      } else {
        r[q].x = (gk < K && gi < M) ? p[0] : E(0);
        r[q].y = (gk < K && gi + 1 < M) ? p[1] : E(0);
      }
    }
  }
}

template <class C, bool AK>
__device__ __forceinline__ void store_a(const typename C::v2 (&r)[C::ACH], typename C::E* As) {
#pragma unroll
  for (int q = 0; q < C::ACH; ++q) {
    const int v = threadIdx.x + 256 * q;
    if (AK) {
      const int row = v / (C::BK / 2), kc = (v % (C::BK / 2)) * 2;
      *reinterpret_cast<typename C::v2*>(As + row * C::PA + kc) = r[q];
    } else {
      const int k = v / (C::BM / 2), row = (v % (C::BM / 2)) * 2;
      As[row * C::PA + k] = r[q].x;
      As[(row + 1) * C::PA + k] = r[q].y;
    }
  }
}

template <class C, bool BNC>
__device__ __forceinline__ void load_b(typename C::v2 (&r)[C::BCH],
                                       const typename C::E* __restrict__ B, int64_t ldb, int N,
                                       int K, int n0, int k0, bool vec) {
  typedef typename C::E E;
  typedef typename C::v2 v2;
#pragma unroll
  for (int q = 0; q < C::BCH; ++q) {
    const int v = threadIdx.x + 256 * q;
    int gj, gk;
    if (BNC) {
      gk = k0 + v / (C::BN / 2);
      gj = n0 + (v % (C::BN / 2)) * 2;
      const E* p = B + (int64_t)gk * ldb + gj;
      if (vec && gk < K && gj + 1 < N) {
        r[q] = *reinterpret_cast<const v2*>(p);
      } else {
        r[q].x = (gk < K && gj < N) ? p[0] : E(0);
        r[q].y = (gk < K && gj + 1 < N) ? p[1] : E(0);
      }
    } else {
      gj = n0 + v / (C::BK / 2);
      gk = k0 + (v % (C::BK / 2)) * 2;
      const E* p = B + gk + (int64_t)gj * ldb;
      if (vec && gj < N && gk + 1 < K) {
        r[q] = *reinterpret_cast<const v2*>(p);
      } else {
        r[q].x = (gj < N && gk < K) ? p[0] : E(0);
        r[q].y = (gj < N && gk + 1 < K) ? p[1] : E(0);
      }
    }
  }
}

template <class C, bool BNC>
__device__ __forceinline__ void store_b(const typename C::v2 (&r)[C::BCH], typename C::E* Bs) {
#pragma unroll
  for (int q = 0; q < C::BCH; ++q) {
    const int v = threadIdx.x + 256 * q;
    if (BNC) {
      const int k = v / (C::BN / 2), col = (v % (C::BN / 2)) * 2;
      *reinterpret_cast<typename C::v2*>(Bs + k * C::PB + col) = r[q];
    } else {
      const int col = v / (C::BK / 2), kc = (v % (C::BK / 2)) * 2;
      Bs[kc * C::PB + col] = r[q].x;
      Bs[(kc + 1) * C::PB + col] = r[q].y;
    }
  }
}

}  // namespace

namespace {

template <class C, bool AK, bool BNC, bool BETA0, int TRI>
__global__ __launch_bounds__(256, 2) void gemm_kernel(int M, int N, int K, typename C::E alpha,
                                                      const typename C::E* __restrict__ A,
                                                      int64_t lda,
                                                      const typename C::E* __restrict__ B,
                                                      int64_t ldb, typename C::E beta,
                                                      typename C::E* __restrict__ Cm, int64_t ldc,
                                                      int tiles_m, int tiles_n, bool veca,
                                                      bool vecb, int kchunk, int64_t a_zstep,
                                                      int64_t b_zstep, int64_t c_zstep) {
  typedef typename C::E E;
  typedef typename C::v4 v4;
  __shared__ __attribute__((aligned(16))) E As[2][C::AS];
  __shared__ __attribute__((aligned(16))) E Bs[2][C::BS];
  // split K: workgroup row z of the grid takes K range [z kchunk, (z + 1) kchunk) and writes
  // its own partial product (the operands and C advanced by z steps)
  {
    const int z = blockIdx.y;
    K = min(kchunk, K - z * kchunk);
    A += z * a_zstep;
    B += z * b_zstep;
    Cm += z * c_zstep;
  }

  int tmi, tni;
  static_assert(TRI == 0 || C::BM == C::BN, "triangular launches need square tiles");
  if (TRI == 1)  // launched over the tiles of the lower triangle only
    lower_tile_of(blockIdx.x, tiles_m, tiles_n, tmi, tni);
  else if (TRI == 2)  // upper: the lower tiles of the transposed grid
    lower_tile_of(blockIdx.x, tiles_n, tiles_m, tni, tmi);
  else
    gemm_tile_of(blockIdx.x, tiles_m, tiles_n, tmi, tni);
  const int m0 = tmi * C::BM, n0 = tni * C::BN;
  if (TRI == 1 && n0 > m0 + C::BM - 1) return;  // tile strictly above the diagonal
  if (TRI == 2 && m0 > n0 + C::BN - 1) return;  // tile strictly below the diagonal

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int l16 = lane & 15, lk = lane >> 4;

  v4 acc[C::TM][C::TN];
#pragma unroll
  for (int i = 0; i < C::TM; ++i)
#pragma unroll
    for (int j = 0; j < C::TN; ++j) acc[i][j] = v4{E(0), E(0), E(0), E(0)};

  typename C::v2 ra[C::ACH], rb[C::BCH];
  const int nk = (K + C::BK - 1) / C::BK;
  load_a<C, AK>(ra, A, lda, M, K, m0, 0, veca);
  load_b<C, BNC>(rb, B, ldb, N, K, n0, 0, vecb);
  store_a<C, AK>(ra, As[0]);
  store_b<C, BNC>(rb, Bs[0]);
  __syncthreads();

  // one K step of the matrix cores on the LDS slices `cur`
  auto kstep = [&](int cur) {
    const E* as = As[cur] + (wm * C::WM + l16) * C::PA + lk;
    const E* bs = Bs[cur] + lk * C::PB + wn * C::WN + l16;
#pragma unroll
    for (int kk = 0; kk < C::BK / 4; ++kk) {
      E a[C::TM], b[C::TN];
#pragma unroll
      for (int i = 0; i < C::TM; ++i) a[i] = as[i * 16 * C::PA + kk * 4];
#pragma unroll
      for (int j = 0; j < C::TN; ++j) b[j] = bs[kk * 4 * C::PB + j * 16];
#pragma unroll
      for (int i = 0; i < C::TM; ++i)
#pragma unroll
        for (int j = 0; j < C::TN; ++j) acc[i][j] = Vec<E>::mfma(a[i], b[j], acc[i][j]);
    }
  };
  for (int kt = 0; kt + 1 < nk; ++kt) {
    const int cur = kt & 1;
    load_a<C, AK>(ra, A, lda, M, K, m0, (kt + 1) * C::BK, veca);
    load_b<C, BNC>(rb, B, ldb, N, K, n0, (kt + 1) * C::BK, vecb);
    kstep(cur);
    store_a<C, AK>(ra, As[cur ^ 1]);
    store_b<C, BNC>(rb, Bs[cur ^ 1]);
    __syncthreads();
  }
  // Tiles wholly inside C (and inside the stored triangle) with beta != 0 (the trailing updates
  // of the factorizations, where K is one block and the read of C is as long as the K loop):
  // the C values of the first PF 16-row groups are fetched while the matrix cores run the last
  // K step, the others in one batch per group below.
  const bool full = m0 + C::BM <= M && n0 + C::BN <= N &&
                    (TRI == 0 || (TRI == 1 && n0 + C::BN - 1 <= m0) ||
                     (TRI == 2 && m0 + C::BM - 1 <= n0));
  constexpr int PF = C::TM < 2 ? C::TM : 2;
  E cp[PF][C::TN][4];
  const int64_t crow0 = (int64_t)(m0 + wm * C::WM);
  const int ccol0 = n0 + wn * C::WN + l16;
  if (!BETA0 && full) {
#pragma unroll
    for (int i = 0; i < PF; ++i)
#pragma unroll
      for (int j = 0; j < C::TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          cp[i][j][r] = Cm[(crow0 + i * 16 + Vec<E>::row(lk, r)) * ldc + ccol0 + j * 16];
  }
  if (nk > 0) kstep((nk - 1) & 1);

  if (!BETA0 && full) {
#pragma unroll
    for (int i = 0; i < C::TM; ++i) {
      E cv[C::TN][4];
#pragma unroll
      for (int j = 0; j < C::TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          cv[j][r] = i < PF ? cp[i < PF ? i : 0][j][r]
                            : Cm[(crow0 + i * 16 + Vec<E>::row(lk, r)) * ldc + ccol0 + j * 16];
#pragma unroll
      for (int j = 0; j < C::TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          Cm[(crow0 + i * 16 + Vec<E>::row(lk, r)) * ldc + ccol0 + j * 16] =
              fma(alpha, acc[i][j][r], beta * cv[j][r]);
    }
    return;
  }

  // epilogue: register r of lane l holds row Vec<E>::row(l/16, r), column l%16 of each tile
#pragma unroll
  for (int i = 0; i < C::TM; ++i)
#pragma unroll
    for (int j = 0; j < C::TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = m0 + wm * C::WM + i * 16 + Vec<E>::row(lk, r);
        const int gj = n0 + wn * C::WN + j * 16 + l16;
        bool ok = gi < M && gj < N;
        if (TRI == 1) ok = ok && gj <= gi;
        if (TRI == 2) ok = ok && gj >= gi;
        if (ok) {
          E* p = Cm + (int64_t)gi * ldc + gj;
          const E v = acc[i][j][r];
          *p = BETA0 ? alpha * v : fma(alpha, v, beta * *p);
        }
      }
}

template <class C, bool AK, bool BNC>
int launch_cfg(int M, int N, int K, typename C::E alpha, const typename C::E* A, int64_t lda,
               const typename C::E* B, int64_t ldb, typename C::E beta, typename C::E* Cm,
               int64_t ldc, int tri, hipStream_t s, int splits = 1, int kchunk = 0) {
  if (splits <= 1) kchunk = K;
  const int64_t az = AK ? (int64_t)kchunk : (int64_t)kchunk * lda;
  const int64_t bz = BNC ? (int64_t)kchunk * ldb : (int64_t)kchunk;
  const int64_t cz = splits > 1 ? (int64_t)M * ldc : 0;
  const int tm = ceil_div(M, C::BM), tn = ceil_div(N, C::BN);
  const int64_t tiles = tri == 1   ? lower_tile_count(tm, tn)
                        : tri == 2 ? lower_tile_count(tn, tm)
                                   : (int64_t)tm * tn;
  if ((int64_t)tm * tn > INT32_MAX) {
    set_error("gemm: problem too large (%lld tiles)", (long long)tm * tn);
    return DDLA_ERR_INVALID_ARG;
  }
  const bool veca = (reinterpret_cast<uintptr_t>(A) % (2 * sizeof(typename C::E)) == 0) && (lda % 2 == 0);
  const bool vecb = (reinterpret_cast<uintptr_t>(B) % (2 * sizeof(typename C::E)) == 0) && (ldb % 2 == 0);
  dim3 grid(static_cast<unsigned>(tiles), static_cast<unsigned>(splits)), block(256);
#define DDLA_GEMM_GO(B0, T)                                                                 \
  hipLaunchKernelGGL((gemm_kernel<C, AK, BNC, B0, T>), grid, block, 0, s, M, N, K, alpha, A, \
                     lda, B, ldb, beta, Cm, ldc, tm, tn, veca, vecb, kchunk, az, bz, cz)
  const bool b0 = (beta == typename C::E(0));
  if (tri == 0) {
    if (b0) DDLA_GEMM_GO(true, 0); else DDLA_GEMM_GO(false, 0);
  } else if (tri == 1) {
    if (b0) DDLA_GEMM_GO(true, 1); else DDLA_GEMM_GO(false, 1);
  } else {
    if (b0) DDLA_GEMM_GO(true, 2); else DDLA_GEMM_GO(false, 2);
  }
#undef DDLA_GEMM_GO
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

template <class C>
int launch_layout(int M, int N, int K, typename C::E alpha, const typename C::E* A, int64_t lda,
                  bool ak, const typename C::E* B, int64_t ldb, bool bn, typename C::E beta,
                  typename C::E* Cm, int64_t ldc, int tri, hipStream_t s, int splits = 1,
                  int kchunk = 0) {
  if (ak && bn)
    return launch_cfg<C, true, true>(M, N, K, alpha, A, lda, B, ldb, beta, Cm, ldc, tri, s,
                                     splits, kchunk);
  if (ak && !bn)
    return launch_cfg<C, true, false>(M, N, K, alpha, A, lda, B, ldb, beta, Cm, ldc, tri, s,
                                      splits, kchunk);
  if (!ak && bn)
    return launch_cfg<C, false, true>(M, N, K, alpha, A, lda, B, ldb, beta, Cm, ldc, tri, s,
                                      splits, kchunk);
  return launch_cfg<C, false, false>(M, N, K, alpha, A, lda, B, ldb, beta, Cm, ldc, tri, s,
                                     splits, kchunk);
}

// C = alpha sum_z P[z] + beta C over the partial products of a split-K GEMM (z in order, so
// the result does not depend on the schedule); tri: 0 all, 1 lower, 2 upper triangle of C
template <typename T>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(int M, int N, int S,
                                                            const T* __restrict__ P, T alpha,
                                                            T beta, T* __restrict__ Cm,
                                                            int64_t ldc, int tri) {
  const int64_t mn = (int64_t)M * N;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < mn;
       e += (int64_t)gridDim.x * 256) {
    const int i = static_cast<int>(e / N), j = static_cast<int>(e % N);
    if ((tri == 1 && j > i) || (tri == 2 && j < i)) continue;
    T acc = P[e];
    for (int z = 1; z < S; ++z) acc += P[z * mn + e];
    T* c = Cm + (int64_t)i * ldc + j;
    *c = beta == T(0) ? alpha * acc : fma(alpha, acc, beta * *c);
  }
}

// C = beta*C over a strided m x n view (tri: 0 all, 1 lower, 2 upper); beta == 0 stores 0.
template <typename T>
__global__ void scale2d_kernel(int64_t m, int64_t n, T beta, T* c, int64_t rs, int64_t cs,
                               int tri) {
  const int64_t total = m * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / n, j = t % n;
    if (tri == 1 && j > i) continue;
    if (tri == 2 && j < i) continue;
    T* p = c + i * rs + j * cs;
    *p = (beta == T(0)) ? T(0) : beta * *p;
  }
}

__global__ void dcopy2d_kernel(int64_t m, int64_t n, const double* __restrict__ src, int64_t rs,
                               int64_t cs, double* __restrict__ dst, int64_t ld) {
  const int64_t total = m * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / n, j = t % n;
    dst[i * ld + j] = src[i * rs + j * cs];
  }
}

__global__ void dscatter2d_kernel(int64_t m, int64_t n, const double* __restrict__ src, int64_t ld,
                                  double* __restrict__ dst, int64_t rs, int64_t cs) {
  const int64_t total = m * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / n, j = t % n;
    dst[i * rs + j * cs] = src[i * ld + j];
  }
}

int grid_for(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return static_cast<int>(g);
}

// FP64 C = alpha A B + beta C for problems with few tiles and a long K (C1's 1024^3: 256
// tiles of 64 x 64 on 256 CUs, one wave per SIMD): row-major A (k contiguous) and B (n
// contiguous), every tile inside C and K a multiple of 32, so the loop body has no bounds
// checks or branches and the scheduler can interleave it.  Within a K step of 32 the k values
// are paired: MFMA pair p (two 16x16x4 products) takes k = 8p + 2 lk + {0, 1} in lane group lk,
// so one 16-byte LDS read gives a lane the operands of both products (A as [m][k], B stored
// transposed as [n][k]); the pairs are summed in the order p = 0..3, k = 8p + 2lk then + 1
// (a fixed order, not the plain k order).  The step runs as 8 groups of 4 independent
// products with the LDS reads of the next pair and the next step's global loads placed between
// groups (sched_barrier fences keep that order), so one wave keeps its matrix core fed.
// Measured at 1024^3: 51 us against 57.7 us for the general 64 x 64 kernel.  (Accumulators must
// stay in VGPRs: __launch_bounds__(256, 2) caps the registers; FP64 MFMAs accumulating in AGPRs
// issue at less than half the rate on gfx950.)
//
// KS = 2 (a grid of no more tiles than CUs, as at 1024^3): 512 threads per tile, waves 0-3
// take the first half of K and waves 4-7 the second, each half with its own LDS buffers, so
// every SIMD runs two waves; at the end the second half's accumulators pass through LDS and
// C = alpha (P0 + P1) (+ beta C), the halves always summed in that order.
template <bool BETA0, int KS>
__global__ __launch_bounds__(256 * KS, KS == 1 ? 2 : 1) void dgemm_long_k_kernel(
    int M, int N, int K, double alpha, const double* __restrict__ A, int64_t lda,
    const double* __restrict__ B, int64_t ldb, double beta, double* __restrict__ Cm,
    int64_t ldc, int tiles_m, int tiles_n) {
  typedef double v2 __attribute__((ext_vector_type(2)));
  typedef double v4 __attribute__((ext_vector_type(4)));
  constexpr int BM = 64, BN = 64, BK = 32, P = BK + 2;  // [row][k] slices, rows of BK + 2
  constexpr int LQ = BM * BK / 2 / 256;         // 16-byte loads per thread and operand
  static_assert(KS == 1 || KS == 2, "one or two K halves");
  __shared__ __attribute__((aligned(16))) double Ash[KS][2][BM * P];
  __shared__ __attribute__((aligned(16))) double Bsh[KS][2][BN * P];
  const int half = KS == 1 ? 0 : (int)(threadIdx.x >> 8);
  double (&As)[2][BM * P] = Ash[half];
  double (&Bs)[2][BN * P] = Bsh[half];
  int tmi, tni;
  gemm_tile_of(blockIdx.x, tiles_m, tiles_n, tmi, tni);
  const int m0 = tmi * BM, n0 = tni * BN;
  const int tid = threadIdx.x & 255, lane = tid & 63, wave = tid >> 6;
  const int nk = K / BK / KS;                 // K steps of each half
  A += (int64_t)half * nk * BK;               // this half's K range
  B += (int64_t)half * nk * BK * ldb;
  const int wm = wave >> 1, wn = wave & 1;
  const int l16 = lane & 15, lk = lane >> 4;
  // loaders: A row ar, k pair ak (v2 along k); B row bk (a k), n pair bn (v2 along n)
  const double* ag[LQ];
  const double* bg[LQ];
  int aoff[LQ], boff[LQ][2];
#pragma unroll
  for (int q = 0; q < LQ; ++q) {
    const int v = tid + 256 * q;
    const int ar = v / (BK / 2), ak = (v % (BK / 2)) * 2;
    ag[q] = A + (int64_t)(m0 + ar) * lda + ak;
    aoff[q] = ar * P + ak;
    const int bk = v / (BN / 2), bn = (v % (BN / 2)) * 2;
    bg[q] = B + (int64_t)bk * ldb + n0 + bn;
    boff[q][0] = bn * P + bk;
    boff[q][1] = (bn + 1) * P + bk;
  }
  v2 ra[LQ], rb[LQ];
  auto gload = [&](int kt) {
#pragma unroll
    for (int q = 0; q < LQ; ++q) {
      ra[q] = *reinterpret_cast<const v2*>(ag[q] + (int64_t)kt * BK);
      rb[q] = *reinterpret_cast<const v2*>(bg[q] + (int64_t)kt * BK * ldb);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < LQ; ++q) {
      *reinterpret_cast<v2*>(&As[buf][aoff[q]]) = ra[q];
      Bs[buf][boff[q][0]] = rb[q].x;
      Bs[buf][boff[q][1]] = rb[q].y;
    }
  };
  v4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4{0.0, 0.0, 0.0, 0.0};
  gload(0);
  lstore(0);
  __syncthreads();
  // One K step as 8 groups of 4 independent products (the 2 x 2 accumulators, so no product
  // waits on the one before it); sched_barrier(0) keeps every group, and the reads placed
  // between groups, in this order: the operands of pair p + 1 are read from LDS while pair p
  // runs, and the next step's global loads (kt + 1 < nk) go out one per group.
  v2 a[2][2], b[2][2];  // [buffer][i or j]
  auto lds_pair = [&](int cur, int p, int buf) {
    const double* as = &As[cur][(wm * 32 + l16) * P + 2 * lk + 8 * p];
    const double* bs = &Bs[cur][(wn * 32 + l16) * P + 2 * lk + 8 * p];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[buf][i] = *reinterpret_cast<const v2*>(as + i * 16 * P);
#pragma unroll
    for (int j = 0; j < 2; ++j) b[buf][j] = *reinterpret_cast<const v2*>(bs + j * 16 * P);
  };
  auto group = [&](int buf, int e) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[buf][i][e], b[buf][j][e], acc[i][j], 0,
                                                         0, 0);
    __builtin_amdgcn_sched_barrier(0);
  };
  // pair 0 of slice `cur` must already be in buffer 0
  auto kstep = [&](int cur, int knext) {
#pragma unroll
    for (int p = 0; p < BK / 8; ++p) {
      const int buf = p & 1;
      if (p + 1 < BK / 8) lds_pair(cur, p + 1, buf ^ 1);
      __builtin_amdgcn_sched_barrier(0);
      group(buf, 0);
      if (knext >= 0 && p < LQ) {
        ra[p] = *reinterpret_cast<const v2*>(ag[p] + (int64_t)knext * BK);
        rb[p] = *reinterpret_cast<const v2*>(bg[p] + (int64_t)knext * BK * ldb);
      }
      __builtin_amdgcn_sched_barrier(0);
      group(buf, 1);
    }
  };
  lds_pair(0, 0, 0);
  for (int kt = 0; kt + 1 < nk; ++kt) {
    const int cur = kt & 1;
    kstep(cur, kt + 1);
    lstore(cur ^ 1);
    __syncthreads();
    lds_pair(cur ^ 1, 0, 0);
  }
  kstep((nk - 1) & 1, -1);
  if constexpr (KS == 2) {
    // the second half's 16 accumulators per thread to LDS (the first half's buffers, free
    // now), the first half adds them to its own: P0 + P1
    __syncthreads();
    double* xch = &Ash[0][0][0];
    if (half == 1) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) xch[((i * 2 + j) * 4 + r) * 256 + tid] = acc[i][j][r];
    }
    __syncthreads();
    if (half == 1) return;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[i][j][r] += xch[((i * 2 + j) * 4 + r) * 256 + tid];
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = m0 + wm * 32 + i * 16 + lk + 4 * r;
        const int gj = n0 + wn * 32 + j * 16 + l16;
        double* p = Cm + (int64_t)gi * ldc + gj;
        *p = BETA0 ? alpha * acc[i][j][r] : fma(alpha, acc[i][j][r], beta * *p);
      }
}

template <typename T>
int gemm_any(int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, bool a_kcontig,
             const T* B, int64_t ldb, bool b_ncontig, T beta, T* C, int64_t ldc, int tri,
             hipStream_t s) {
  if (m <= 0 || n <= 0) return 0;
  if (k <= 0 || alpha == T(0)) {
    if (beta == T(1)) return 0;
    hipLaunchKernelGGL(scale2d_kernel<T>, dim3(grid_for(m * n)), dim3(256), 0, s, m, n, beta, C,
                       ldc, (int64_t)1, tri);
    DDLA_HIP_CHECK(hipGetLastError());
    return 0;
  }
  if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) {
    set_error("gemm: dimension exceeds 2^31-1");
    return DDLA_ERR_INVALID_ARG;
  }
  const int M = (int)m, N = (int)n, K = (int)k;
  const int64_t big_tiles = (int64_t)ceil_div(M, 128) * ceil_div(N, 128);
  const int cus = ctx().num_cus > 0 ? ctx().num_cus : 256;
  const double mn = (tri ? 0.5 : 1.0) * (double)m * n;
  ProfScope prof(sizeof(T) == 8 ? "dgemm_mfma_f64" : "sgemm_mfma_f32", 2.0 * mn * k,
                 sizeof(T) * ((double)m * k + (double)k * n + (beta == T(0) ? 1.0 : 2.0) * mn), s);
  if (big_tiles >= 2 * cus)
    return launch_layout<Cfg<T, 128, 128>>(M, N, K, alpha, A, lda, a_kcontig, B, ldb, b_ncontig,
                                           beta, C, ldc, tri, s);
  // few output tiles and a long K (products like U^T C of tall panels): split K over a second
  // grid dimension, partial products in a workspace, then one ordered reduction.  Only on the
  // library's main stream (one workspace).
  const int64_t tiles64 = (int64_t)ceil_div(M, 64) * ceil_div(N, 64);
  if (s == stream() && tiles64 * 2 < cus && K >= 1024) {
    int splits = (int)std::min<int64_t>(K / 512, ceil_div(2 * cus, tiles64));
    if (splits >= 2) {
      const int kchunk = (ceil_div(K, splits) + 31) / 32 * 32;
      splits = ceil_div(K, kchunk);
      void* ws = nullptr;
      DDLA_RETURN_IF(workspace(WS_SPLITK, (size_t)splits * M * N * sizeof(T), &ws));
      T* P = static_cast<T*>(ws);
      DDLA_RETURN_IF((launch_layout<Cfg<T, 64, 64, 32>>(M, N, K, T(1), A, lda, a_kcontig, B, ldb,
                                                        b_ncontig, T(0), P, N, 0, s, splits,
                                                        kchunk)));
      hipLaunchKernelGGL(splitk_reduce_kernel<T>, dim3(grid_for((int64_t)M * N)), dim3(256), 0, s,
                         M, N, splits, P, alpha, beta, C, ldc, tri);
      DDLA_HIP_CHECK(hipGetLastError());
      return 0;
    }
  }
  // small tiles: one wave per SIMD, so each K step is made long enough (32) for the MFMA work
  // to cover the latency of the next step's global loads
  if constexpr (sizeof(T) == 8) {
    if (tri == 0 && a_kcontig && b_ncontig && M % 64 == 0 && N % 64 == 0 && K % 32 == 0 &&
        K >= 512 && lda % 2 == 0 && ldb % 2 == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0 &&
        (reinterpret_cast<uintptr_t>(B) & 15) == 0) {
      const int tm = M / 64, tn = N / 64;
      // no more tiles than CUs: two K halves per tile, so each SIMD runs two waves
      const bool two = (int64_t)tm * tn <= cus && K % 64 == 0;
#define DLK(B0, KS)                                                                          \
  hipLaunchKernelGGL((dgemm_long_k_kernel<B0, KS>), dim3(tm * tn), dim3(256 * KS), 0, s, M, N, \
                     K, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn)
      if (beta == T(0)) {
        if (two) DLK(true, 2); else DLK(true, 1);
      } else {
        if (two) DLK(false, 2); else DLK(false, 1);
      }
#undef DLK
      DDLA_HIP_CHECK(hipGetLastError());
      return 0;
    }
  }
  return launch_layout<Cfg<T, 64, 64, 32>>(M, N, K, alpha, A, lda, a_kcontig, B, ldb, b_ncontig,
                                           beta, C, ldc, tri, s);
}

}  // namespace

int launch_dscale2d(int64_t m, int64_t n, double beta, double* c, int64_t rs, int64_t cs, int tri,
                    hipStream_t s) {
  if (m <= 0 || n <= 0 || beta == 1.0) return 0;
  hipLaunchKernelGGL(scale2d_kernel<double>, dim3(grid_for(m * n)), dim3(256), 0, s, m, n, beta, c,
                     rs, cs, tri);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_dcopy2d(int64_t m, int64_t n, const double* src, int64_t rs, int64_t cs, double* dst,
                   int64_t ld, hipStream_t s) {
  if (m <= 0 || n <= 0) return 0;
  hipLaunchKernelGGL(dcopy2d_kernel, dim3(grid_for(m * n)), dim3(256), 0, s, m, n, src, rs, cs,
                     dst, ld);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_dscatter2d(int64_t m, int64_t n, const double* src, int64_t ld, double* dst, int64_t rs,
                      int64_t cs, hipStream_t s) {
  if (m <= 0 || n <= 0) return 0;
  hipLaunchKernelGGL(dscatter2d_kernel, dim3(grid_for(m * n)), dim3(256), 0, s, m, n, src, ld, dst,
                     rs, cs);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_dgemm(int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                 bool a_kcontig, const double* B, int64_t ldb, bool b_ncontig, double beta,
                 double* C, int64_t ldc, int tri, hipStream_t s) {
  return gemm_any<double>(m, n, k, alpha, A, lda, a_kcontig, B, ldb, b_ncontig, beta, C, ldc, tri, s);
}

int launch_sgemm(int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                 bool a_kcontig, const float* B, int64_t ldb, bool b_ncontig, float beta, float* C,
                 int64_t ldc, int tri, hipStream_t s) {
  return gemm_any<float>(m, n, k, alpha, A, lda, a_kcontig, B, ldb, b_ncontig, beta, C, ldc, tri, s);
}

}  // namespace ddla
