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
//      smallest value, exactly AMax's rule) and publishes (key, row, candidate row
//      contents) in a double-buffered slot, then raises its step flag;
//   2. every workgroup waits for all flags, reduces the W candidates to the same winner,
//      reads the winner's row from its slot, and performs the interchange locally (the
//      owner of the diagonal row also published that row, so the owner of the pivot row
//      can take it);
//   3. every workgroup scales its part of the column by 1/pivot (InvScal multiplies by
//      the reciprocal, DDLA/BLAS/1v.cs:102) and applies the rank-1 update to its rows.
// Every row carries one extra LDS column: the panel-relative index of the row whose
// trailing data belongs there.  At the end each workgroup emits (dst, src) pairs for
// moved rows, so the interchanges for the rest of the matrix are one parallel gather
// (laswp_list_kernel) instead of nb sequential swaps.
#include "ddla_internal.hpp"

#include <climits>

namespace ddla {

struct PanelSync {
  // double-buffered exchange area, laid out in one workspace allocation
  unsigned* flags;     // [2][W]
  double* keys;        // [2][W]
  int* rows;           // [2][W]
  void* cand;          // [2][W][NBX] elements of T
  void* diag;          // [2][NBX]
  int* list;           // [2*nb][2] (dst, src) panel-relative
  int* list_count;     // [1]
};

namespace {

constexpr int NB_MAX = 128;
constexpr int NBX = NB_MAX + 1;  // + origin column

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

template <typename T>
__device__ __forceinline__ T ld_agent(const T* p) {
  if constexpr (sizeof(T) == 8) {
    unsigned long long v = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p),
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_bit_cast(T, v);
  } else {
    unsigned v = __hip_atomic_load(reinterpret_cast<const unsigned*>(p), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_bit_cast(T, v);
  }
}

template <typename T>
__device__ __forceinline__ void st_agent(T* p, T v) {
  if constexpr (sizeof(T) == 8) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                       __builtin_bit_cast(unsigned long long, v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  } else {
    __hip_atomic_store(reinterpret_cast<unsigned*>(p), __builtin_bit_cast(unsigned, v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
  unsigned* flags = sy.flags;
  double* keys = sy.keys;
  int* rowsv = sy.rows;
  T* cand = static_cast<T*>(sy.cand);
  T* diag = static_cast<T*>(sy.diag);

  if (w == 0 && tid == 0) st_agent(sy.list_count, 0);
  // an earlier panel already hit a zero pivot: leave everything as it is (empty list)
  if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;

  for (int idx = tid; idx < nloc * nb; idx += 256) {
    const int r = idx / nb, c = idx % nb;
    tile[r * NBX + c] = P[(int64_t)(r0 + r) * ld + c];
  }
  for (int r = tid; r < nloc; r += 256) tile[r * NBX + NB_MAX] = static_cast<T>(r0 + r);
  __syncthreads();

  int stop = 0;
  for (int i = 0; i < nb; ++i) {
    const int par = i & 1;
    const unsigned seq = base + static_cast<unsigned>(i) + 1u;
    // 1. local candidate
    double key = -2.0;
    int row = INT_MAX;
    if (tid < nloc && r0 + tid >= i) {
      const double v = fabs(static_cast<double>(tile[tid * NBX + i]));
      key = (v != v) ? -1.0 : v;
      row = r0 + tid;
    }
    block_argmax(key, row, sk, sr);
    // 2. publish candidate row, diagonal row, key; raise flag
    T* my_cand = cand + ((size_t)par * W + w) * NBX;
    if (key > -2.0) {
      const int lr = row - r0;
      for (int c = tid; c < NBX; c += 256) st_agent(my_cand + c, tile[lr * NBX + c]);
    }
    const bool own_i = (i >= r0 && i < r0 + nloc);
    if (own_i) {
      for (int c = tid; c < NBX; c += 256) {
        const T v = tile[(i - r0) * NBX + c];
        st_agent(diag + (size_t)par * NBX + c, v);
        drow[c] = v;
      }
    }
    if (tid == 0) {
      st_agent(keys + (size_t)par * W + w, key);
      __hip_atomic_store(rowsv + (size_t)par * W + w, row, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
    __syncthreads();
    if (tid == 0)
      __hip_atomic_store(flags + (size_t)par * W + w, seq, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
    // 3. wait for every workgroup's flag of this step
    if (wave == 0) {
      for (;;) {
        bool ok = true;
        for (int q = lane; q < W; q += 64)
          ok = ok && (__hip_atomic_load(flags + (size_t)par * W + q, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT) == seq);
        if (__all(ok)) break;
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    // 4. global winner (identical in every workgroup)
    key = -2.0;
    row = INT_MAX;
    for (int q = tid; q < W; q += 256) {
      const double k2 = ld_agent(keys + (size_t)par * W + q);
      const int r2 = __hip_atomic_load(rowsv + (size_t)par * W + q, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
      better(key, row, k2, r2);
    }
    block_argmax(key, row, sk, sr);
    const int wwin = row / R;
    for (int c = tid; c < NBX; c += 256)
      prow[c] = ld_agent(cand + ((size_t)par * W + wwin) * NBX + c);
    __syncthreads();
    const T pv = prow[i];
    if (w == 0 && tid == 0) piv[i] = row - i;
    if (pv == static_cast<T>(0)) {
      if (w == 0 && tid == 0) atomicCAS(info, 0, info_base + i + 1);
      stop = 1;
    }
    // 5. interchange rows i and `row`
    if (row != i) {
      if (own_i)
        for (int c = tid; c < NBX; c += 256) tile[(i - r0) * NBX + c] = prow[c];
      if (row >= r0 && row < r0 + nloc) {
        for (int c = tid; c < NBX; c += 256)
          tile[(row - r0) * NBX + c] = own_i ? drow[c] : ld_agent(diag + (size_t)par * NBX + c);
      }
    }
    __syncthreads();
    if (stop) break;
    // 6. scale the column below the pivot and update the rest of the panel rows
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

}  // namespace

// exchange area for panel kernels (sizes for up to 1024 workgroups)
int panel_sync(PanelSync* sy, int nb, unsigned** base_counter) {
  constexpr int WMAX = 1024;
  const size_t bytes = 2 * WMAX * sizeof(unsigned) + 2 * WMAX * sizeof(double) +
                       2 * WMAX * sizeof(int) + 2 * (size_t)WMAX * NBX * sizeof(double) +
                       2 * NBX * sizeof(double) + 4 * (size_t)NB_MAX * sizeof(int) + 256;
  void* ws = nullptr;
  static bool zeroed = false;
  static void* zeroed_ptr = nullptr;
  DDLA_RETURN_IF(workspace(WS_PANEL, bytes, &ws));
  if (!zeroed || zeroed_ptr != ws) {
    DDLA_HIP_CHECK(hipMemset(ws, 0, bytes));
    zeroed = true;
    zeroed_ptr = ws;
  }
  char* p = static_cast<char*>(ws);
  sy->flags = reinterpret_cast<unsigned*>(p);
  p += 2 * WMAX * sizeof(unsigned);
  sy->keys = reinterpret_cast<double*>(p);
  p += 2 * WMAX * sizeof(double);
  sy->rows = reinterpret_cast<int*>(p);
  p += 2 * WMAX * sizeof(int);
  sy->cand = p;
  p += 2 * (size_t)WMAX * NBX * sizeof(double);
  sy->diag = p;
  p += 2 * NBX * sizeof(double);
  sy->list = reinterpret_cast<int*>(p);
  p += 4 * (size_t)NB_MAX * sizeof(int);
  sy->list_count = reinterpret_cast<int*>(p);
  (void)nb;
  static unsigned counter = 0;
  *base_counter = &counter;
  return 0;
}

template <typename T>
int launch_getf2_panel_ex(int64_t m, int nb, T* P, int64_t ld, int32_t* piv, int32_t* info,
                          int info_base, int* list_out, int* count_out, hipStream_t s) {
  if (nb > NB_MAX || nb <= 0) {
    set_error("panel width %d outside 1..%d", nb, NB_MAX);
    return DDLA_ERR_INVALID_ARG;
  }
  PanelSync sy;
  unsigned* counter = nullptr;
  DDLA_RETURN_IF(panel_sync(&sy, nb, &counter));
  if (list_out) {
    sy.list = list_out;
    sy.list_count = count_out;
  }
  const unsigned base = *counter;
  *counter += static_cast<unsigned>(nb) + 2u;
  const int W32 = ceil_div(m, 32), W64 = ceil_div(m, 64), W128 = ceil_div(m, 128);
  ProfScope prof(sizeof(T) == 8 ? "getf2_panel_f64" : "getf2_panel_f32",
                 (double)m * nb * nb - (double)nb * nb * nb / 3.0, 2.0 * sizeof(T) * m * nb, s);
  int rc;
  if (W32 <= 128 && W32 <= max_coresident<T, 32>())
    rc = launch_panel_r<T, 32>((int)m, nb, P, ld, piv, info, sy, base, info_base, s);
  else if (W64 <= 256 && W64 <= max_coresident<T, 64>())
    rc = launch_panel_r<T, 64>((int)m, nb, P, ld, piv, info, sy, base, info_base, s);
  else if (W128 <= 1024 && W128 <= max_coresident<T, 128>())
    rc = launch_panel_r<T, 128>((int)m, nb, P, ld, piv, info, sy, base, info_base, s);
  else {
    set_error("panel of %lld rows exceeds the co-resident capacity", (long long)m);
    return DDLA_ERR_UNSUPPORTED;
  }
  return rc;
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
