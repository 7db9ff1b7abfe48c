// Kernels of the blocked LU with partial pivoting (FP64 reference path; the FP32 panel of
// the mixed-precision path reuses them through the element type T).
//
// Reference: PLUDecBlock / PLUDecUnblock / GetAndApplyPivUnblock / ApplyPivInternel
// (DDLA/Factorizations/LU.cs:289-429), AMax (DDLA/BLAS/1v.cs:30-41).
//
// Panel (getf2_panel_kernel): the m x nb panel is spread over W co-resident workgroups,
// R rows each, held in LDS for the whole panel.  Partial pivoting needs one global
// arg-max per column; each column step is one exchange round between the workgroups:
//   1. every workgroup reduces |a_ri| over its rows (first index wins ties, NaN is the
//      smallest value, exactly AMax's rule) and publishes its candidate (key, row) and the
//      candidate row's contents; the owner of the diagonal row also publishes that row;
//   2. every workgroup reads the W candidates, reduces them to the same winner, reads the
//      winner's row, and performs the interchange locally (the owner of the pivot row takes
//      the published diagonal row);
//   3. every workgroup scales its part of the column by 1/pivot (InvScal multiplies by
//      the reciprocal, DDLA/BLAS/1v.cs:102) and applies the rank-1 update to its rows.
// The exchange needs no fence or barrier flag: every published 64-bit word carries the
// 32-bit step tag next to 32 bits of payload (an FP64 value takes two words), is written
// and read with single-copy-atomic agent-scope accesses, and a reader simply re-reads a
// word until its tag is the current step's.  Slots alternate between two buffers by step
// parity; a workgroup can only reach step i+2 after every workgroup published step i+1,
// i.e. after every reader of step i finished, so a slot is never overwritten while read.
// Every row carries one extra LDS column: the panel-relative index of the row whose
// trailing data belongs there.  At the end each workgroup emits (dst, src) pairs for
// moved rows, so the interchanges for the rest of the matrix are one parallel gather
// (laswp_list_kernel) instead of nb sequential swaps.
#include "ddla_internal.hpp"

#include <climits>

namespace ddla {

struct PanelSync {
  // double-buffered exchange area (one workspace allocation), 64-bit tagged words
  unsigned long long* keys;  // [2][WMAX][3]: key high, key low, row
  unsigned long long* cand;  // [2][WMAX][NBX][2]: candidate row (2 words per element)
  unsigned long long* diag;  // [2][NBX][2]: diagonal row of the step
  int* list;                 // [2*nb][2] (dst, src) panel-relative
  int* list_count;           // [1]
};

namespace {

constexpr int NB_MAX = 128;
constexpr int NBX = NB_MAX + 1;  // + origin column
constexpr int WMAX = 256;        // one exchange slot per lane of a workgroup

typedef unsigned long long u64;

__device__ __forceinline__ void better(double& k, int& r, double k2, int r2) {
  if (k2 > k || (k2 == k && r2 < r)) {
    k = k2;
    r = r2;
  }
}

// all 256 lanes end with the same (key, row)
__device__ __forceinline__ void block_argmax(double& key, int& row, double* sk, int* sr) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double k2 = __shfl_xor(key, off);
    const int r2 = __shfl_xor(row, off);
    better(key, row, k2, r2);
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) {
    sk[wave] = key;
    sr[wave] = row;
  }
  __syncthreads();
  key = sk[0];
  row = sr[0];
#pragma unroll
  for (int w = 1; w < 4; ++w) better(key, row, sk[w], sr[w]);
}

__device__ __forceinline__ void put(u64* p, unsigned payload, unsigned tag) {
  __hip_atomic_store(p, (static_cast<u64>(payload) << 32) | tag, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 get(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// waits for the word to carry `tag`; returns its payload
__device__ __forceinline__ unsigned wait_word(const u64* p, unsigned tag) {
  u64 v = get(p);
  while (static_cast<unsigned>(v) != tag) {
    __builtin_amdgcn_s_sleep(1);
    v = get(p);
  }
  return static_cast<unsigned>(v >> 32);
}

template <typename T>
__device__ __forceinline__ void put_elem(u64* p, T v, unsigned tag) {
  if constexpr (sizeof(T) == 8) {
    const u64 b = __builtin_bit_cast(u64, v);
    put(p, static_cast<unsigned>(b >> 32), tag);
    put(p + 1, static_cast<unsigned>(b), tag);
  } else {
    put(p, __builtin_bit_cast(unsigned, v), tag);
  }
}
template <typename T>
__device__ __forceinline__ T wait_elem(const u64* p, unsigned tag) {
  if constexpr (sizeof(T) == 8) {
    const u64 hi = wait_word(p, tag);
    const u64 lo = wait_word(p + 1, tag);
    return __builtin_bit_cast(T, (hi << 32) | lo);
  } else {
    return __builtin_bit_cast(T, wait_word(p, tag));
  }
}

template <typename T, int R>
__global__ __launch_bounds__(256) void getf2_panel_kernel(int m, int nb, T* __restrict__ P,
                                                          int64_t ld, int32_t* __restrict__ piv,
                                                          int32_t* __restrict__ info,
                                                          PanelSync sy, unsigned base,
                                                          int info_base) {
  __shared__ T tile[R * NBX + 2 * NBX];
  __shared__ double sk[4];
  __shared__ int sr[4];
  T* prow = tile + R * NBX;
  T* drow = prow + NBX;

  const int W = gridDim.x;
  const int w = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r0 = w * R;
  const int nloc = max(0, min(R, m - r0));

  if (w == 0 && tid == 0)
    __hip_atomic_store(sy.list_count, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // an earlier panel already hit a zero pivot: leave everything as it is (empty list)
  if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;

  for (int idx = tid; idx < nloc * nb; idx += 256) {
    const int r = idx / nb, c = idx % nb;
    tile[r * NBX + c] = P[(int64_t)(r0 + r) * ld + c];
  }
  for (int r = tid; r < nloc; r += 256) tile[r * NBX + NB_MAX] = static_cast<T>(r0 + r);
  __syncthreads();

  for (int i = 0; i < nb; ++i) {
    const int par = i & 1;
    const unsigned tag = base + static_cast<unsigned>(i) + 1u;
    // 1. local candidate, published with the contents of its row
    double key = -2.0;
    int row = INT_MAX;
    if (tid < nloc && r0 + tid >= i) {
      const double v = fabs(static_cast<double>(tile[tid * NBX + i]));
      key = (v != v) ? -1.0 : v;
      row = r0 + tid;
    }
    block_argmax(key, row, sk, sr);
    if (key > -2.0) {
      u64* my = sy.cand + ((size_t)par * WMAX + w) * NBX * 2;
      const T* src = tile + (row - r0) * NBX;
      for (int c = tid; c < NBX; c += 256) put_elem<T>(my + 2 * c, src[c], tag);
    }
    const bool own_i = (i >= r0 && i < r0 + nloc);
    if (own_i) {
      u64* dg = sy.diag + (size_t)par * NBX * 2;
      for (int c = tid; c < NBX; c += 256) {
        const T v = tile[(i - r0) * NBX + c];
        put_elem<T>(dg + 2 * c, v, tag);
        drow[c] = v;
      }
    }
    if (tid == 0) {
      u64* kp = sy.keys + ((size_t)par * WMAX + w) * 3;
      const u64 kb = __builtin_bit_cast(u64, key);
      put(kp, static_cast<unsigned>(kb >> 32), tag);
      put(kp + 1, static_cast<unsigned>(kb), tag);
      put(kp + 2, static_cast<unsigned>(row), tag);
    }
    // 2. global winner (identical in every workgroup): lane q reads slot q
    key = -2.0;
    row = INT_MAX;
    if (tid < W) {
      const u64* kp = sy.keys + ((size_t)par * WMAX + tid) * 3;
      const u64 hi = wait_word(kp, tag);
      const u64 lo = wait_word(kp + 1, tag);
      key = __builtin_bit_cast(double, (hi << 32) | lo);
      row = static_cast<int>(wait_word(kp + 2, tag));
    }
    block_argmax(key, row, sk, sr);
    const int wwin = row / R;
    const bool own_p = (row != i) && (row >= r0 && row < r0 + nloc);
    {
      const u64* cw = sy.cand + ((size_t)par * WMAX + wwin) * NBX * 2;
      for (int c = tid; c < NBX; c += 256) prow[c] = wait_elem<T>(cw + 2 * c, tag);
      if (own_p && !own_i) {
        const u64* dg = sy.diag + (size_t)par * NBX * 2;
        for (int c = tid; c < NBX; c += 256) drow[c] = wait_elem<T>(dg + 2 * c, tag);
      }
    }
    __syncthreads();
    const T pv = prow[i];
    if (w == 0 && tid == 0) piv[i] = row - i;
    const bool stop = (pv == static_cast<T>(0));
    if (stop && w == 0 && tid == 0) atomicCAS(info, 0, info_base + i + 1);
    // 3. interchange rows i and `row`
    if (row != i) {
      if (own_i)
        for (int c = tid; c < NBX; c += 256) tile[(i - r0) * NBX + c] = prow[c];
      if (own_p)
        for (int c = tid; c < NBX; c += 256) tile[(row - r0) * NBX + c] = drow[c];
    }
    __syncthreads();
    if (stop) break;
    // 4. scale the column below the pivot and update the rest of the panel rows
    const T rinv = static_cast<T>(1) / pv;
    for (int lr = wave; lr < nloc; lr += 4) {
      if (r0 + lr <= i) continue;
      T* tr = tile + lr * NBX;
      const T l = tr[i] * rinv;
      for (int c = i + 1 + lane; c < nb; c += 64) tr[c] = fma(-l, prow[c], tr[c]);
      if (lane == 0) tr[i] = l;
    }
    __syncthreads();
  }

  // write back and emit the row movement list (dst <- src, panel-relative)
  for (int idx = tid; idx < nloc * nb; idx += 256) {
    const int r = idx / nb, c = idx % nb;
    P[(int64_t)(r0 + r) * ld + c] = tile[r * NBX + c];
  }
  for (int r = tid; r < nloc; r += 256) {
    const int src = static_cast<int>(tile[r * NBX + NB_MAX]);
    if (src != r0 + r) {
      const int slot = atomicAdd(sy.list_count, 1);
      sy.list[2 * slot] = r0 + r;
      sy.list[2 * slot + 1] = src;
    }
  }
}

// Applies the (dst <- src) row list of a panel to columns [0, ncols) of A (rows relative to
// row0), skipping columns [skip_lo, skip_hi).  Two phases per column through tmp so that a
// source row is read before any destination write.
template <typename T>
__global__ void laswp_list_kernel(T* __restrict__ A, int64_t ld, int64_t row0, int64_t ncols,
                                  int64_t skip_lo, int64_t skip_hi, const int* __restrict__ list,
                                  const int* __restrict__ count, T* __restrict__ tmp) {
  const int cnt = *count;
  const int64_t width = ncols - (skip_hi - skip_lo);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < width;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = (t < skip_lo) ? t : t + (skip_hi - skip_lo);
    for (int e = 0; e < cnt; ++e) tmp[(int64_t)e * width + t] = A[(row0 + list[2 * e + 1]) * ld + c];
    for (int e = 0; e < cnt; ++e) A[(row0 + list[2 * e]) * ld + c] = tmp[(int64_t)e * width + t];
  }
}

// perm[i] = i, then the panel lists are composed in order: after panel k,
// perm[row0_k + dst] = perm_prev[row0_k + src].  One workgroup, perm in LDS.
__global__ __launch_bounds__(256) void perm_compose_kernel(int n, int npanels, int nb,
                                                           const int* __restrict__ lists,
                                                           const int* __restrict__ counts,
                                                           int* __restrict__ perm_out) {
  extern __shared__ int perm[];
  for (int i = threadIdx.x; i < n; i += 256) perm[i] = i;
  __syncthreads();
  for (int k = 0; k < npanels; ++k) {
    const int row0 = k * nb;
    const int cnt = counts[k];
    const int* L = lists + (size_t)k * 4 * nb;
    int v[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = threadIdx.x + 256 * q;
      v[q] = (e < cnt) ? perm[row0 + L[2 * e + 1]] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = threadIdx.x + 256 * q;
      if (e < cnt) perm[row0 + L[2 * e]] = v[q];
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += 256) perm_out[i] = perm[i];
}

// Sequential application of relative pivots to an index vector (one lane, LDS).
__global__ void piv_to_perm_kernel(int n, const int32_t* __restrict__ piv, int* __restrict__ out) {
  extern __shared__ int perm[];
  for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < n; ++i) {
      const int p = piv[i];
      if (p > 0) {
        const int a = perm[i], b = perm[i + p];
        perm[i] = b;
        perm[i + p] = a;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) out[i] = perm[i];
}

// dst[i] = src[perm[i]] (forward) or dst[perm[i]] = src[i] (inverse), rows of ncols.
template <typename T>
__global__ void gather_rows_kernel(int64_t n, int64_t ncols, const T* __restrict__ src,
                                   int64_t lds, const int* __restrict__ perm, T* __restrict__ dst,
                                   int64_t ldd, bool inverse) {
  const int64_t total = n * ncols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / ncols, c = t % ncols;
    if (!inverse)
      dst[i * ldd + c] = src[(int64_t)perm[i] * lds + c];
    else
      dst[(int64_t)perm[i] * ldd + c] = src[i * lds + c];
  }
}

int grid_for(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return static_cast<int>(g);
}

template <typename T, int R>
int launch_panel_r(int m, int nb, T* P, int64_t ld, int32_t* piv, int32_t* info,
                   const PanelSync& sy, unsigned base, int info_base, hipStream_t s) {
  const int W = ceil_div(m, R);
  void* args[] = {&m, &nb, &P, &ld, &piv, &info, const_cast<PanelSync*>(&sy), &base, &info_base};
  DDLA_HIP_CHECK(hipLaunchCooperativeKernel(
      reinterpret_cast<const void*>(getf2_panel_kernel<T, R>), dim3(W), dim3(256), args, 0, s));
  return 0;
}

template <typename T, int R>
int max_coresident() {
  static int cached = -1;
  if (cached < 0) {
    int per_cu = 0;
    DDLA_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(getf2_panel_kernel<T, R>), 256, 0));
    cached = per_cu * ctx().num_cus;
  }
  return cached;
}

template <typename T, int R>
bool fits(int64_t m, int limit) {
  const int64_t W = (m + R - 1) / R;
  return W <= limit && W <= max_coresident<T, R>();
}

}  // namespace

// exchange area for panel kernels
int panel_sync(PanelSync* sy, unsigned** base_counter) {
  const size_t kb = 2 * (size_t)WMAX * 3 * sizeof(u64);
  const size_t cb = 2 * (size_t)WMAX * NBX * 2 * sizeof(u64);
  const size_t db = 2 * (size_t)NBX * 2 * sizeof(u64);
  const size_t lb = 4 * (size_t)NB_MAX * sizeof(int) + 256;
  const size_t bytes = kb + cb + db + lb;
  void* ws = nullptr;
  static void* zeroed_ptr = nullptr;
  DDLA_RETURN_IF(workspace(WS_PANEL, bytes, &ws));
  if (zeroed_ptr != ws) {  // tags start from a clean slate once per allocation
    DDLA_HIP_CHECK(hipMemsetAsync(ws, 0, bytes, stream()));
    zeroed_ptr = ws;
  }
  char* p = static_cast<char*>(ws);
  sy->keys = reinterpret_cast<u64*>(p);
  sy->cand = reinterpret_cast<u64*>(p + kb);
  sy->diag = reinterpret_cast<u64*>(p + kb + cb);
  sy->list = reinterpret_cast<int*>(p + kb + cb + db);
  sy->list_count = sy->list + 4 * NB_MAX;
  static unsigned counter = 0;
  *base_counter = &counter;
  return 0;
}

template <typename T>
int launch_getf2_panel_ex(int64_t m, int nb, T* P, int64_t ld, int32_t* piv, int32_t* info,
                          int info_base, int* list_out, int* count_out, hipStream_t s) {
  if (nb > NB_MAX || nb <= 0 || m < nb) {
    set_error("panel %lld x %d outside 1 <= nb <= min(m, %d)", (long long)m, nb, NB_MAX);
    return DDLA_ERR_INVALID_ARG;
  }
  PanelSync sy;
  unsigned* counter = nullptr;
  DDLA_RETURN_IF(panel_sync(&sy, &counter));
  if (list_out) {
    sy.list = list_out;
    sy.list_count = count_out;
  }
  // step tags: every launch takes a fresh range, so stale words never match
  const unsigned base = *counter;
  *counter += static_cast<unsigned>(nb) + 2u;
  if (*counter < base) *counter = 1;  // wrapped: restart (all words hold older tags)
  ProfScope prof(sizeof(T) == 8 ? "getf2_panel_f64" : "getf2_panel_f32",
                 (double)m * nb * nb - (double)nb * nb * nb / 3.0, 2.0 * sizeof(T) * m * nb, s);
  // fewest workgroups that hold the panel (fewer slots to read per step)
  if (fits<T, 32>(m, 64)) return launch_panel_r<T, 32>((int)m, nb, P, ld, piv, info, sy, base, info_base, s);
  if (fits<T, 64>(m, 128)) return launch_panel_r<T, 64>((int)m, nb, P, ld, piv, info, sy, base, info_base, s);
  if (fits<T, 128>(m, WMAX)) return launch_panel_r<T, 128>((int)m, nb, P, ld, piv, info, sy, base, info_base, s);
  if constexpr (sizeof(T) == 4) {
    if (fits<T, 256>(m, WMAX)) return launch_panel_r<T, 256>((int)m, nb, P, ld, piv, info, sy, base, info_base, s);
  }
  set_error("panel of %lld rows exceeds the co-resident capacity", (long long)m);
  return DDLA_ERR_UNSUPPORTED;
}

template <typename T>
int launch_laswp_list(T* A, int64_t ld, int64_t row0, int64_t ncols, int64_t skip_lo,
                      int64_t skip_hi, const int* list, const int* count, int max_count,
                      hipStream_t s) {
  const int64_t width = ncols - (skip_hi - skip_lo);
  if (width <= 0) return 0;
  void* tmp = nullptr;
  DDLA_RETURN_IF(workspace(WS_TEMP, (size_t)max_count * width * sizeof(T), &tmp));
  // bytes: every moved row (at most max_count) is read and written once
  ProfScope prof("laswp_list", 0.0, 2.0 * sizeof(T) * max_count * width, s);
  hipLaunchKernelGGL(laswp_list_kernel<T>, dim3(grid_for(width)), dim3(256), 0, s, A, ld, row0,
                     ncols, skip_lo, skip_hi, list, count, static_cast<T*>(tmp));
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_perm_compose(int n, int npanels, int nb, const int* lists, const int* counts, int* perm,
                        hipStream_t s) {
  const size_t sh = (size_t)n * sizeof(int);
  if (sh > 160 * 1024) {
    set_error("permutation of %d rows does not fit in LDS", n);
    return DDLA_ERR_UNSUPPORTED;
  }
  static bool attr = false;
  if (!attr) {
    DDLA_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(perm_compose_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr = true;
  }
  hipLaunchKernelGGL(perm_compose_kernel, dim3(1), dim3(256), sh, s, n, npanels, nb, lists, counts,
                     perm);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_piv_to_perm(int64_t n, const int32_t* piv, int32_t* perm, hipStream_t s) {
  const size_t sh = (size_t)n * sizeof(int);
  if (sh > 160 * 1024) {
    set_error("permutation of %lld rows does not fit in LDS", (long long)n);
    return DDLA_ERR_UNSUPPORTED;
  }
  static bool attr = false;
  if (!attr) {
    DDLA_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(piv_to_perm_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr = true;
  }
  hipLaunchKernelGGL(piv_to_perm_kernel, dim3(1), dim3(256), sh, s, (int)n, piv, perm);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

template <typename T>
int launch_gather_rows(int64_t n, int64_t ncols, const T* src, int64_t lds, const int32_t* perm,
                       T* dst, int64_t ldd, bool inverse, hipStream_t s) {
  if (n <= 0 || ncols <= 0) return 0;
  hipLaunchKernelGGL(gather_rows_kernel<T>, dim3(grid_for(n * ncols)), dim3(256), 0, s, n, ncols,
                     src, lds, perm, dst, ldd, inverse);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

template int launch_getf2_panel_ex<double>(int64_t, int, double*, int64_t, int32_t*, int32_t*,
                                           int, int*, int*, hipStream_t);
template int launch_getf2_panel_ex<float>(int64_t, int, float*, int64_t, int32_t*, int32_t*, int,
                                          int*, int*, hipStream_t);
template int launch_laswp_list<double>(double*, int64_t, int64_t, int64_t, int64_t, int64_t,
                                       const int*, const int*, int, hipStream_t);
template int launch_laswp_list<float>(float*, int64_t, int64_t, int64_t, int64_t, int64_t,
                                      const int*, const int*, int, hipStream_t);
template int launch_gather_rows<double>(int64_t, int64_t, const double*, int64_t, const int32_t*,
                                        double*, int64_t, bool, hipStream_t);
template int launch_gather_rows<float>(int64_t, int64_t, const float*, int64_t, const int32_t*,
                                       float*, int64_t, bool, hipStream_t);

}  // namespace ddla
