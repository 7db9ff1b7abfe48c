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

constexpr int BS_INV = 64;  // order of the inverted diagonal blocks
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

// Bound of a wait for another workgroup's word: past `limit` poll rounds the waiter stores
// the call's tag in the abort word and gives up; a waiter that finds the tag there (another
// workgroup gave up) gives up too.  The tag is never 0, so a cleared word names no call.
struct WaitBound {
  unsigned* abort;
  unsigned limit;
  unsigned tag;
  __device__ __forceinline__ bool keep_waiting(unsigned& rounds) const {
    ++rounds;
    if (rounds > limit) {
      __hip_atomic_store(abort, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    return (rounds & 63u) != 0 ||
           __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != tag;
  }
};

// the value published at p for this call; `failed` is set (and 0 returned) when the wait gives up
template <typename T>
__device__ __forceinline__ T await(const u64* p, const WaitBound& wb, bool& failed) {
  const unsigned tag = wb.tag;
  unsigned rounds = 0;
  if constexpr (sizeof(T) == 8) {
    u64 a = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    u64 b = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (static_cast<unsigned>(a) != tag || static_cast<unsigned>(b) != tag) {
      if (!wb.keep_waiting(rounds)) {
        failed = true;
        return T(0);
      }
      __builtin_amdgcn_s_sleep(1);
      a = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      b = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return __builtin_bit_cast(T, ((a >> 32) << 32) | (b >> 32));
  } else {
    u64 a = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (static_cast<unsigned>(a) != tag) {
      if (!wb.keep_waiting(rounds)) {
        failed = true;
        return T(0);
      }
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
  constexpr int BS = BS_INV;
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

// block rows per step of the solve: 128 for FP32 (64 KB blocks), 64 for FP64
template <typename T>
constexpr int solve_bs() {
  return sizeof(T) == 4 ? 128 : 64;
}

// Layout of the prepared data behind Dinv (trsv_prepare): the inverted 64 x 64 diagonal
// blocks, then per block row kb of the solve (BS rows) Dfull_kb = inverse of the whole
// diagonal block, then M_kb = Dfull_kb * op(T)[kb, kprev], kprev = the block solved just
// before kb, then N_kb = Dfull_kb * op(T)[kb, kprev2], kprev2 = the one before kprev.  With
// them x_kb = Dfull_kb (b_kb - sum_{older blocks} T x) - N_kb x_kprev2 - M_kb x_kprev: each
// of the two latest neighbours enters through a single product.
__host__ __device__ inline int64_t blocks_of(int64_t n, int64_t b) { return (n + b - 1) / b; }
__host__ __device__ inline int64_t inv64_elems(int64_t n) {
  return blocks_of(n, BS_INV) * BS_INV * BS_INV;
}
template <int BS>
__host__ __device__ inline int64_t blocks_elems(int64_t n) { return blocks_of(n, BS) * BS * BS; }

// out (BS x BS, row-major) = S1 S2 for two BS x BS LDS blocks.  Thread (tr, tc) =
// (tid / 16, tid % 16) owns rows tr + 16 i and columns tc + 16 j (an RB x RB register tile),
// so that a k step reads RB values of each operand for RB * RB fmas; every output is still
// one fma chain over k = 0, 1, ..., BS - 1 from zero.
template <typename T, int BS>
__device__ __forceinline__ void block_product(const T (*S1)[BS + 1], const T (*S2)[BS + 1],
                                              T* __restrict__ out) {
  constexpr int RB = BS / 16;
  const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
  T acc[RB][RB];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < RB; ++j) acc[i][j] = T(0);
  for (int k = 0; k < BS; ++k) {
    T a[RB], b[RB];
#pragma unroll
    for (int i = 0; i < RB; ++i) a[i] = S1[tr + 16 * i][k];
#pragma unroll
    for (int j = 0; j < RB; ++j) b[j] = S2[k][tc + 16 * j];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
      for (int j = 0; j < RB; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
  }
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < RB; ++j) out[(tr + 16 * i) * BS + tc + 16 * j] = acc[i][j];
}

// Builds Dfull and M (blockIdx.y == 0) or N (blockIdx.y == 1), see above, for one block row
// per workgroup, in LDS.  Both halves build Dfull; the first stores it.
template <typename T, int BS, bool LOWER, bool TRANS>
__global__ __launch_bounds__(256) void trsv_combine_kernel(int64_t n, const T* __restrict__ A,
                                                           int64_t lda, T* __restrict__ Dinv) {
  constexpr int NSUB = BS / BS_INV;
  __shared__ T S1[BS][BS + 1];  // Dfull
  __shared__ T S2[BS][BS + 1];  // operand blocks
  const int nblk = (int)blocks_of(n, BS);
  const int kb = blockIdx.x;
  const bool second = blockIdx.y == 1;
  {
    const int kneed = second ? (LOWER ? kb - 2 : kb + 2) : kb;
    if (kneed < 0 || kneed >= nblk) return;  // uniform: no N block for this row
  }
  const int64_t r0 = (int64_t)kb * BS;
  const int tid = threadIdx.x;
  T* Dfull = Dinv + inv64_elems(n) + (int64_t)kb * BS * BS;
  T* Mb = Dinv + inv64_elems(n) + blocks_elems<BS>(n) + (int64_t)kb * BS * BS;
  const int nb64 = (int)blocks_of(n, BS_INV);
  auto opT = [&](int64_t r, int64_t c) -> T {  // element (r, c) of op(T), zero outside n
    if (r >= n || c >= n) return T(0);
    return TRANS ? A[c * lda + r] : A[r * lda + c];
  };
  // inverted 64-blocks on the diagonal (identity beyond n)
  for (int idx = tid; idx < BS * BS; idx += 256) {
    const int r = idx / BS, c = idx % BS;
    const int sr = r / BS_INV, sc = c / BS_INV;
    T v = T(0);
    if (sr == sc) {
      const int64_t b64 = (int64_t)kb * NSUB + sr;
      v = b64 < nb64 ? Dinv[b64 * BS_INV * BS_INV + (r % BS_INV) * BS_INV + (c % BS_INV)]
                     : T(r == c ? 1 : 0);
    }
    S1[r][c] = v;
  }
  if (NSUB == 2) {
    // LOWER: X = -D1 L10 D0 (rows 64.., cols ..64); UPPER: Y = -D0 U01 D1 (rows ..64, cols 64..)
    constexpr int H = BS_INV;
    for (int idx = tid; idx < H * H; idx += 256) {
      const int r = idx / H, c = idx % H;
      S2[r][c] = LOWER ? opT(r0 + H + r, r0 + c) : opT(r0 + r, r0 + H + c);
    }
    __syncthreads();
    // P = LOWER ? L10 D0 : U01 D1  -> S2[.][H..2H)
    for (int idx = tid; idx < H * H; idx += 256) {
      const int r = idx / H, c = idx % H;
      T acc = T(0);
      for (int k = 0; k < H; ++k) acc = fma(S2[r][k], LOWER ? S1[k][c] : S1[H + k][H + c], acc);
      S2[r][H + c] = acc;
    }
    __syncthreads();
    // off-diagonal quarter = -(LOWER ? D1 P : D0 P)
    for (int idx = tid; idx < H * H; idx += 256) {
      const int r = idx / H, c = idx % H;
      T acc = T(0);
      for (int k = 0; k < H; ++k) acc = fma(LOWER ? S1[H + r][H + k] : S1[r][k], S2[k][H + c], acc);
      if (LOWER) S1[H + r][c] = -acc;
      else S1[r][H + c] = -acc;
    }
  }
  __syncthreads();
  if (!second)
    for (int idx = tid; idx < BS * BS; idx += 256) Dfull[idx] = S1[idx / BS][idx % BS];
  // M: the block solved just before kb; N: the one before that
  const int kprev = second ? (LOWER ? kb - 2 : kb + 2) : (LOWER ? kb - 1 : kb + 1);
  if (kprev < 0 || kprev >= nblk) return;  // uniform
  const int64_t c0 = (int64_t)kprev * BS;
  __syncthreads();  // S2 is rewritten
  for (int idx = tid; idx < BS * BS; idx += 256) {
    const int r = idx / BS, c = idx % BS;
    S2[r][c] = opT(r0 + r, c0 + c);
  }
  __syncthreads();
  block_product<T, BS>(S1, S2, second ? Mb + blocks_elems<BS>(n) : Mb);
}

template <typename T, int BS, int NR, bool LOWER, bool TRANS>
__global__ __launch_bounds__(256) void trsv_kernel(int64_t n, int nrhs, const T* __restrict__ A,
                                                   int64_t lda, const T* __restrict__ Dinv,
                                                   T* __restrict__ B, int64_t ldb,
                                                   u64* __restrict__ XW, unsigned tag,
                                                   int split, const int* __restrict__ skip,
                                                   unsigned* __restrict__ abort_word,
                                                   unsigned limit) {
  // a refinement that has already converged queues its remaining steps as no-ops: every
  // workgroup reads the same word (written by an earlier kernel) and leaves before any wait
  if (skip && *skip) return;
  const WaitBound wb{abort_word, limit, tag};
  bool failed = false;  // a wait of this thread gave up (the workgroup leaves at the next check)
  constexpr int Q = 256 / BS;       // threads per row of a block product (k split)
  constexpr int KQ = BS / Q;        // k range per thread
  constexpr int E = BS * BS / 256;  // staged block elements per thread
  constexpr int XPT = (BS * NR + 255) / 256;  // X / B words per thread
  constexpr int NW = Words<T>::N;
  constexpr int VW = 16 / sizeof(T);  // elements per 16-byte load
  constexpr int EV = E / VW;
  typedef T vecT __attribute__((ext_vector_type(VW)));
  const bool avec = lda % VW == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0;
  __shared__ T Ts[BS][BS + 1];
  __shared__ T Ms[BS][BS + 1];
  __shared__ T xs[BS][NR];
  __shared__ T ys[BS][NR];
  __shared__ T red[Q][BS][NR];
  const int nblk = (int)((n + BS - 1) / BS);
  const int G = gridDim.x;
  const int tid = threadIdx.x;
  const int row = tid % BS, q = tid / BS;
  const T* Dfull = Dinv + inv64_elems(n);
  const T* Mall = Dfull + blocks_elems<BS>(n);
  const T* Nall = Mall + blocks_elems<BS>(n);

  // split (one block row per workgroup, 2 nblk workgroups): workgroup nblk + jb helps block
  // row jb by streaming its odd blocks and publishing the partial sums (HW words)
  const bool helper = split && (int)blockIdx.x >= nblk;
  u64* HW = XW + n * nrhs * 2;
  for (int jb = helper ? (int)blockIdx.x - nblk : (int)blockIdx.x; jb < nblk;
       jb += split ? nblk : G) {
    const int kb = LOWER ? jb : nblk - 1 - jb;  // dependency order
    const int64_t r0 = (int64_t)kb * BS;
    const int rows = (int)min<int64_t>(BS, n - r0);
    // Blocks in dependency order jc < jb - 1 stream through Ts (block jc + 1 and the X it
    // multiplies are loaded while block jc is used, so a long block row runs at the memory
    // system's pace), then Dfull; the block kprev = jb - 1 enters through M.
    T stage[E];
    // Staging: element e of vector v of this thread is (i, j) = (idx / (BS/VW), (idx % (BS/VW))
    // VW + e) with idx = tid + 256 v, j running along memory; 16-byte loads wherever the
    // row is aligned (few wide loads keep the address units from limiting the stream).
    auto load_block = [&](int jc) {
      const int kc = LOWER ? jc : nblk - 1 - jc;
      const int64_t c0 = (int64_t)kc * BS;
      const int cols = (int)min<int64_t>(BS, n - c0);
#pragma unroll
      for (int v = 0; v < EV; ++v) {
        const int idx = tid + 256 * v;
        const int i = idx / (BS / VW), j0 = (idx % (BS / VW)) * VW;
        // !TRANS: op(T)(r0 + i, c0 + j) = A[(r0 + i) lda + c0 + j]
        //  TRANS: op(T)(r0 + j, c0 + i) = A[(c0 + i) lda + r0 + j]
        const bool rowok = TRANS ? (i < cols) : (i < rows);
        const int lim = TRANS ? rows : cols;
        const T* p = TRANS ? A + (c0 + i) * lda + r0 + j0 : A + (r0 + i) * lda + c0 + j0;
        if (avec && rowok && j0 + VW <= lim) {
          const vecT x = *reinterpret_cast<const vecT*>(p);
#pragma unroll
          for (int e = 0; e < VW; ++e) stage[v * VW + e] = x[e];
        } else {
#pragma unroll
          for (int e = 0; e < VW; ++e) stage[v * VW + e] = (rowok && j0 + e < lim) ? p[e] : T(0);
        }
      }
    };
    auto load_dense = [&](const T* src) {  // a BS x BS row-major block (16-byte aligned)
#pragma unroll
      for (int v = 0; v < EV; ++v) {
        const vecT x = *reinterpret_cast<const vecT*>(src + (int64_t)(tid + 256 * v) * VW);
#pragma unroll
        for (int e = 0; e < VW; ++e) stage[v * VW + e] = x[e];
      }
    };
    auto load_dfull = [&]() { load_dense(Dfull + (int64_t)kb * BS * BS); };
    auto store_block = [&](bool tr, T (*dst)[BS + 1]) {
#pragma unroll
      for (int v = 0; v < EV; ++v) {
        const int idx = tid + 256 * v;
        const int i = idx / (BS / VW), j0 = (idx % (BS / VW)) * VW;
#pragma unroll
        for (int e = 0; e < VW; ++e) {
          if (tr) dst[j0 + e][i] = stage[v * VW + e];
          else dst[i][j0 + e] = stage[v * VW + e];
        }
      }
    };
    u64 xw[XPT][NW];
    auto load_x = [&](int jc) {
      const int kc = LOWER ? jc : nblk - 1 - jc;
      const int64_t c0 = (int64_t)kc * BS;
#pragma unroll
      for (int t = 0; t < XPT; ++t) {
        const int idx = tid + 256 * t;
        const int k = idx / NR, c = idx % NR;
        if (idx < BS * NR && c < nrhs && c0 + k < n) {
#pragma unroll
          for (int w = 0; w < NW; ++w)
            xw[t][w] = __hip_atomic_load(XW + ((c0 + k) * nrhs + c) * NW + w, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    };
    auto use_x = [&](int jc) {  // xw (re-read until current) -> xs
      const int kc = LOWER ? jc : nblk - 1 - jc;
      const int64_t c0 = (int64_t)kc * BS;
#pragma unroll
      for (int t = 0; t < XPT; ++t) {
        const int idx = tid + 256 * t;
        if (idx >= BS * NR) break;
        const int k = idx / NR, c = idx % NR;
        T v = T(0);
        if (c < nrhs && c0 + k < n) {
          const u64* p = XW + ((c0 + k) * nrhs + c) * NW;
          unsigned rounds = 0;
          for (;;) {
            bool ok = true;
#pragma unroll
            for (int w = 0; w < NW; ++w) ok = ok && static_cast<unsigned>(xw[t][w]) == tag;
            if (ok || failed) break;
            if (!wb.keep_waiting(rounds)) {
              failed = true;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
#pragma unroll
            for (int w = 0; w < NW; ++w)
              xw[t][w] = __hip_atomic_load(p + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
          if constexpr (NW == 2)
            v = __builtin_bit_cast(T, ((xw[t][0] >> 32) << 32) | (xw[t][1] >> 32));
          else
            v = __builtin_bit_cast(T, static_cast<unsigned>(xw[t][0] >> 32));
        }
        xs[k][c] = v;
      }
    };
    // sum over the Q k-ranges: red[q][r][c] -> dst[r][c] (optionally minus)
    auto product = [&](const T (&acc)[NR]) {
#pragma unroll
      for (int c = 0; c < NR; ++c) red[q][row][c] = acc[c];
    };

    // operands that do not depend on X: M (to LDS, through the staging registers), this
    // thread's k-range of N's row (registers), and B
    if (jb > 0) {
      load_dense(Mall + (int64_t)kb * BS * BS);
      store_block(false, Ms);
    }
    T nreg[KQ];
    if (jb > 1) {
      const T* np = Nall + (int64_t)kb * BS * BS + (int64_t)row * BS + q * KQ;
      constexpr int NV = KQ * sizeof(T) / 16;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const vecT x = *reinterpret_cast<const vecT*>(np + v * VW);
#pragma unroll
        for (int e = 0; e < VW; ++e) nreg[v * VW + e] = x[e];
      }
    }
    T breg[XPT];
#pragma unroll
    for (int t = 0; t < XPT; ++t) {
      const int idx = tid + 256 * t;
      const int r = idx / NR, c = idx % NR;
      breg[t] = (idx < BS * NR && c < nrhs && r < rows) ? B[(r0 + r) * ldb + c] : T(0);
    }
    T part[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) part[c] = T(0);
    const int nstream = jb > 1 ? jb - 2 : 0;
    const int jstep = split ? 2 : 1, jfirst = helper ? 1 : 0;
    const bool helped = split && nstream / 2 > 0;  // the helper streams a block of this row
    if (jfirst < nstream) {
      load_x(jfirst);
      load_block(jfirst);
    } else if (!helper) {
      load_dfull();
    }
    for (int jc = jfirst; jc < nstream; jc += jstep) {
      const int jn = jc + jstep;
      __syncthreads();  // the previous block's product is done with Ts / xs
      store_block(TRANS, Ts);
      if (jn < nstream) load_block(jn);  // the next block's loads overlap the wait for x
      else if (!helper) load_dfull();
      use_x(jc);
      if (jn < nstream) load_x(jn);
      if (__syncthreads_or(failed)) return;
#pragma unroll 4
      for (int k = q * KQ; k < q * KQ + KQ; ++k) {
        const T a = Ts[row][k];
#pragma unroll
        for (int c = 0; c < NR; ++c) part[c] = fma(a, xs[k][c], part[c]);
      }
    }
    if (helper) {  // partial sums of the odd blocks to the row's workgroup
      __syncthreads();
      product(part);
      __syncthreads();
      if (nstream / 2 > 0) {
        for (int idx = tid; idx < BS * NR; idx += 256) {
          const int r = idx / NR, c = idx % NR;
          T sum = T(0);
#pragma unroll
          for (int qq = 0; qq < Q; ++qq) sum += red[qq][r][c];
          if (r < rows && c < nrhs) publish<T>(HW + ((r0 + r) * nrhs + c) * NW, sum, tag);
        }
      }
      __syncthreads();
      continue;
    }
    // y = Dfull (b - sum)
    __syncthreads();
    store_block(false, Ts);
    product(part);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < XPT; ++t) {
      const int idx = tid + 256 * t;
      if (idx < BS * NR) {
        const int r = idx / NR, c = idx % NR;
        T sum = T(0);
#pragma unroll
        for (int qq = 0; qq < Q; ++qq) sum += red[qq][r][c];
        if (helped && c < nrhs && r < rows)
          sum += await<T>(HW + ((r0 + r) * nrhs + c) * NW, wb, failed);
        xs[r][c] = (c < nrhs && r < rows) ? breg[t] - sum : T(0);
      }
    }
    if (__syncthreads_or(failed)) return;
    T acc[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[c] = T(0);
#pragma unroll 4
    for (int k = q * KQ; k < q * KQ + KQ; ++k) {
      const T d = Ts[row][k];
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = fma(d, xs[k][c], acc[c]);
    }
    product(acc);
    __syncthreads();
    for (int idx = tid; idx < BS * NR; idx += 256) {
      const int r = idx / NR, c = idx % NR;
      T sum = T(0);
#pragma unroll
      for (int qq = 0; qq < Q; ++qq) sum += red[qq][r][c];
      ys[r][c] = sum;
    }
    // y -= N x_kprev2 (the block solved two steps before)
    if (jb > 1) {
      load_x(jb - 2);  // use_x only re-reads words that do not carry this call's tag yet
      use_x(jb - 2);
      if (__syncthreads_or(failed)) return;
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = T(0);
#pragma unroll
      for (int kk = 0; kk < KQ; ++kk) {
#pragma unroll
        for (int c = 0; c < NR; ++c) acc[c] = fma(nreg[kk], xs[q * KQ + kk][c], acc[c]);
      }
      product(acc);
      __syncthreads();
      for (int idx = tid; idx < BS * NR; idx += 256) {
        const int r = idx / NR, c = idx % NR;
        T sum = T(0);
#pragma unroll
        for (int qq = 0; qq < Q; ++qq) sum += red[qq][r][c];
        ys[r][c] -= sum;
      }
      __syncthreads();  // xs is rewritten
    }
    // x = y - M x_kprev: the only step that waits for the block solved just before
    if (jb > 0) {
      load_x(jb - 1);
      use_x(jb - 1);
      if (__syncthreads_or(failed)) return;
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = T(0);
#pragma unroll 4
      for (int k = q * KQ; k < q * KQ + KQ; ++k) {
        const T m = Ms[row][k];
#pragma unroll
        for (int c = 0; c < NR; ++c) acc[c] = fma(m, xs[k][c], acc[c]);
      }
      product(acc);
      __syncthreads();
      for (int idx = tid; idx < BS * NR; idx += 256) {
        const int r = idx / NR, c = idx % NR;
        T sum = T(0);
#pragma unroll
        for (int qq = 0; qq < Q; ++qq) sum += red[qq][r][c];
        ys[r][c] -= sum;
      }
    }
    __syncthreads();
    for (int idx = tid; idx < BS * NR; idx += 256) {
      const int r = idx / NR, c = idx % NR;
      if (r < rows && c < nrhs) {
        const T x = ys[r][c];
        publish<T>(XW + ((r0 + r) * nrhs + c) * NW, x, tag);
        B[(r0 + r) * ldb + c] = x;
      }
    }
    __syncthreads();  // the LDS is reused by the next block row of this workgroup
  }
}

constexpr size_t TRSV_ABORT_BYTES = 256;

template <typename T, bool LOWER, bool TRANS>
int max_trsv_groups() {
  static int cached = -1;
  if (cached < 0) {
    int per_cu = 0;
    DDLA_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(trsv_kernel<T, solve_bs<T>(), 8, LOWER, TRANS>),
        256, 0));
    cached = per_cu * ctx().num_cus;
  }
  return cached;
}

// B[0:rows, c0:c0+CW] <- D B[0:rows, c0:c0+CW] in place (D: rows x rows of a BS x BS
// row-major block), one workgroup per CW columns: the workgroup stages its slice of B and D in
// LDS before it writes, so the product reads no output.  Thread t computes rows t % BS of
// columns (t / BS) * CW / 2 ... + CW / 2; D's rows are padded in LDS (the lanes of a read
// take one column of D each, in distinct banks).
template <int BS, int CW>
__global__ __launch_bounds__(256) void block_apply_kernel(int rows, int nrhs,
                                                          const float* __restrict__ D,
                                                          float* __restrict__ B, int64_t ldb) {
  static_assert(BS == 128 && CW % 2 == 0, "256 threads: two column halves of BS rows");
  __shared__ float ds[BS][BS + 1];
  __shared__ float xs[BS][CW];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * CW;
  for (int idx = tid; idx < BS * CW; idx += 256) {
    const int r = idx / CW, c = idx % CW;
    xs[r][c] = (r < rows && c0 + c < nrhs) ? B[(int64_t)r * ldb + c0 + c] : 0.f;
  }
  // D is a whole BS x BS block in memory: its reads are unconditional, issued 16 at a time
  constexpr int PER = BS * BS / 256, BATCH = 16;
#pragma unroll
  for (int b0 = 0; b0 < PER; b0 += BATCH) {
    float v[BATCH];
#pragma unroll
    for (int q = 0; q < BATCH; ++q) v[q] = D[(b0 + q) * 256 + tid];
#pragma unroll
    for (int q = 0; q < BATCH; ++q) {
      const int idx = (b0 + q) * 256 + tid, r = idx / BS, k = idx % BS;
      ds[r][k] = (r < rows && k < rows) ? v[q] : 0.f;
    }
  }
  __syncthreads();
  const int r = tid % BS, h = tid / BS;
  float acc[CW / 2];
#pragma unroll
  for (int c = 0; c < CW / 2; ++c) acc[c] = 0.f;
#pragma unroll 8
  for (int k = 0; k < BS; ++k) {
    const float d = ds[r][k];
#pragma unroll
    for (int c = 0; c < CW / 2; ++c) acc[c] = fmaf(d, xs[k][h * (CW / 2) + c], acc[c]);
  }
  if (r < rows) {
#pragma unroll
    for (int c = 0; c < CW / 2; ++c) {
      const int cc = c0 + h * (CW / 2) + c;
      if (cc < nrhs) B[(int64_t)r * ldb + cc] = acc[c];
    }
  }
}

}  // namespace

// Dinv (nblk x 64 x 64) for op(T); nblk = ceil(n/64)
template <typename T>
int trsv_prepare(bool lower, bool unit, bool trans, int64_t n, const T* A, int64_t lda, T* Dinv,
                 hipStream_t s) {
  const int nblk = ceil_div(n, BS_INV);
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
  constexpr int BS = solve_bs<T>();
  const int nbb = ceil_div(n, BS);
#define TC(L, TR) \
  hipLaunchKernelGGL((trsv_combine_kernel<T, BS, L, TR>), dim3(nbb, 2), dim3(256), 0, s, n, A, \
                     lda, Dinv)
  if (lower) {
    if (trans) TC(true, true); else TC(true, false);
  } else {
    if (trans) TC(false, true); else TC(false, false);
  }
#undef TC
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

// op(T) X = B for many right-hand sides (FP32), level-3 form on the prepared blocks: per
// block row in dependency order, X_kb = Dfull_kb B_kb in place (block_apply_kernel), then the
// block rows still to come take B -= op(T)[later, kb] X_kb (one GEMM).  Same (lower, trans)
// convention and prepared data as trsv_solve.
int trsv_solve_blocked(bool lower, bool trans, int64_t n, int nrhs, const float* A, int64_t lda,
                       const float* Dinv, float* B, int64_t ldb, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return 0;
  constexpr int BS = solve_bs<float>();
  const int64_t nblk = blocks_of(n, BS);
  const float* Dfull = Dinv + inv64_elems(n);
  constexpr int CW = 8;  // right-hand-side columns per workgroup of the diagonal-block product
  const dim3 grid((unsigned)ceil_div(nrhs, CW));
  for (int64_t jb = 0; jb < nblk; ++jb) {
    const int64_t kb = lower ? jb : nblk - 1 - jb;
    const int64_t r0 = kb * BS;
    const int rows = (int)std::min<int64_t>(BS, n - r0);
    // X_kb = Dfull_kb B_kb, in place
    hipLaunchKernelGGL((block_apply_kernel<BS, CW>), grid, dim3(256), 0, s, rows, nrhs,
                       Dfull + kb * BS * BS, B + r0 * ldb, ldb);
    DDLA_HIP_CHECK(hipGetLastError());
    // op(T)[r][c] is A[r lda + c], or A[c lda + r] when trans
    const int64_t lo = lower ? r0 + rows : 0, m = lower ? n - r0 - rows : r0;
    if (m <= 0) continue;
    const float* M = trans ? A + r0 * lda + lo : A + lo * lda + r0;
    DDLA_RETURN_IF(launch_sgemm(m, nrhs, rows, -1.f, M, lda, !trans, B + r0 * ldb, ldb, true,
                                1.f, B + lo * ldb, ldb, 0, s));
  }
  return 0;
}

unsigned* trsv_abort_word() { return static_cast<unsigned*>(ctx().ws[WS_GMRES]); }

// inverted 64-blocks + Dfull + M + N (sized for the larger FP32 solve block)
size_t trsv_dinv_elems(int64_t n) {
  return (size_t)(inv64_elems(n) + 3 * blocks_elems<128>(n));
}

// op(T) X = B in place, nrhs <= 8, with Dinv from trsv_prepare (same lower/trans).
template <typename T>
int trsv_solve(bool lower, bool trans, int64_t n, int nrhs, const T* A, int64_t lda, const T* Dinv,
               T* B, int64_t ldb, hipStream_t s, const int* skip) {
  if (n <= 0 || nrhs <= 0) return 0;
  if (nrhs > RMAX) {
    set_error("trsv: nrhs %d > %d", nrhs, RMAX);
    return DDLA_ERR_INVALID_ARG;
  }
  void* ws = nullptr;
  // the abort word (first 256 bytes), then the X words, then the helpers' partial-sum words
  // (2 words per FP64 value)
  const size_t wbytes = TRSV_ABORT_BYTES + (size_t)n * nrhs * 4 * sizeof(u64);
  // one tag sequence for the FP32 and FP64 solves: both use this workspace, so a word left
  // by a solve of either type never carries the tag of a later solve
  unsigned tag = 0;
  DDLA_RETURN_IF(next_tag(WS_GMRES, wbytes, 1, s, &tag));
  DDLA_RETURN_IF(workspace(WS_GMRES, wbytes, &ws));
  const int nblk = ceil_div(n, solve_bs<T>());
  ProfScope prof(sizeof(T) == 8 ? "trsv_f64" : "trsv_f32", (double)n * n * nrhs,
                 (0.5 * n * n + 2.0 * n * nrhs) * sizeof(T), s);
  unsigned* abort_word = static_cast<unsigned*>(ws);
  u64* XW = reinterpret_cast<u64*>(static_cast<char*>(ws) + TRSV_ABORT_BYTES);
  const unsigned limit = ctx().wait_limit;
  // A plain launch of at most the co-resident number of workgroups: block row jb waits only on
  // the rows before it, which lower workgroup ids (dispatched first) or this workgroup's
  // previous rows compute, and on its helper, which is launched only when every workgroup fits
  // at once.  Waits are bounded (the abort word) in case other work keeps workgroups out.
  // A helper workgroup per block row when twice the rows fit co-resident: the longest rows
  // stream their blocks with two workgroups.
#define TS(NR, L, TR)                                                                       \
  do {                                                                                      \
    const int maxg = max_trsv_groups<T, L, TR>();                                           \
    const int split = 2 * nblk <= maxg ? 1 : 0;                                             \
    const int G = split ? 2 * nblk : std::min(nblk, maxg);                                  \
    hipLaunchKernelGGL((trsv_kernel<T, solve_bs<T>(), NR, L, TR>), dim3(G), dim3(256), 0, s, \
                       n, nrhs, A, lda, Dinv, B, ldb, XW, tag, split, skip, abort_word,      \
                       limit);                                                              \
    DDLA_HIP_CHECK(hipGetLastError());                                                      \
  } while (0)
#define TSN(L, TR)                          \
  do {                                      \
    if (nrhs == 1) TS(1, L, TR);            \
    else if (nrhs == 2) TS(2, L, TR);       \
    else if (nrhs <= 4) TS(4, L, TR);       \
    else TS(8, L, TR);                      \
  } while (0)
  if (lower) {
    if (trans) TSN(true, true); else TSN(true, false);
  } else {
    if (trans) TSN(false, true); else TSN(false, false);
  }
#undef TSN
#undef TS
  return 0;
}

template int trsv_prepare<float>(bool, bool, bool, int64_t, const float*, int64_t, float*, hipStream_t);
template int trsv_prepare<double>(bool, bool, bool, int64_t, const double*, int64_t, double*, hipStream_t);
template int trsv_solve<float>(bool, bool, int64_t, int, const float*, int64_t, const float*, float*,
                               int64_t, hipStream_t, const int*);
template int trsv_solve<double>(bool, bool, int64_t, int, const double*, int64_t, const double*,
                                double*, int64_t, hipStream_t, const int*);

}  // namespace ddla
