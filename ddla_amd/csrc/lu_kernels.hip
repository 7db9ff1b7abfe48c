// Kernels of the blocked LU with partial pivoting (FP64 reference path; the FP32 panel of
// the mixed-precision path reuses them through the element type T).
//
// Reference: PLUDecBlock / PLUDecUnblock / GetAndApplyPivUnblock / ApplyPivInternel
// (DDLA/Factorizations/LU.cs:289-429), AMax (DDLA/BLAS/1v.cs:30-41).
//
// Panel (getf2_panel_kernel): the m x nb panel is spread over W co-resident workgroups,
// R rows each, held in registers for the whole panel.  Partial pivoting needs one global
// arg-max per column; each column step is one exchange round between the workgroups:
//   1. every workgroup reduces |a_ri| over its rows that are not yet pivotal (first index
//      wins ties, NaN is the smallest value, exactly AMax's rule) and publishes its candidate
//      (key, label) and the candidate row's contents;
//   2. every workgroup reads the W candidates, reduces them to the same winner and reads the
//      winner's row;
//   3. every workgroup scales its part of the column by 1/pivot (InvScal multiplies by
//      the reciprocal, DDLA/BLAS/1v.cs:102) and applies the rank-1 update to its rows.
// Interchanges move labels, not data: every register row carries its logical row index
// (its position in the reference's physically interchanged panel).  Swapping rows i and p
// gives the pivot row label i and the row labelled i label p; arg-max ties go to the smallest
// label, the rows labelled < i are pivotal and excluded.  The arithmetic on every row is the
// same as with moved rows.  At the end each row is written to the position of its label,
// and each workgroup emits (dst, src) pairs for the rows that moved, so the interchanges for
// the rest of the matrix are one parallel gather and scatter (laswp_list_kernel) instead of
// nb sequential swaps.
// The exchange needs no fence or barrier flag: every published 64-bit word carries the
// 32-bit step tag next to 32 bits of payload (an FP64 value takes two words), is written
// and read with single-copy-atomic agent-scope accesses, and a reader simply re-reads a
// word until its tag is the current step's.  Slots alternate between two buffers by step
// parity; a workgroup can only reach step i+2 after every workgroup published step i+1,
// i.e. after every reader of step i finished, so a slot is never overwritten while read.
#include "ddla_internal.hpp"
#include "wave_ops.hpp"

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>

namespace ddla {

struct PanelSync {
  // double-buffered exchange area (one workspace allocation), 64-bit tagged words
  unsigned long long* keys;  // [2][WMAX][KSTRIDE]: key (FP64: high, low), packed label (+ pad)
  unsigned long long* cand;  // [2][WMAX][NB_MAX][2]: candidate row (2 words per element)
  int* list;                 // [2*nb][2] (dst, src) panel-relative
  int* list_count;           // [1]
};

namespace {

constexpr int NB_MAX = 128;
constexpr int WMAX = 256;        // at most one workgroup per CU
constexpr int KSTRIDE = 16;      // one 128-byte line per key slot: readers spread over lines

typedef unsigned long long u64;

// The exchange area: keys, candidate rows, the row list and its count, then the abort pair
// (word 0: the first tag of a launch that gave up waiting; word 1: the poll-round limit,
// 0 meaning DEFAULT_WAIT_LIMIT).  The kernel finds the abort pair at a fixed offset from the
// keys, so it costs no kernel argument.
constexpr size_t PANEL_KEY_BYTES = 2 * (size_t)WMAX * KSTRIDE * sizeof(u64);
constexpr size_t PANEL_CAND_BYTES = 2 * (size_t)WMAX * NB_MAX * 2 * sizeof(u64);
constexpr size_t PANEL_LIST_BYTES = 4 * (size_t)NB_MAX * sizeof(int) + 256;
constexpr size_t PANEL_ABORT_OFF = PANEL_KEY_BYTES + PANEL_CAND_BYTES + PANEL_LIST_BYTES;
constexpr size_t PANEL_ABORT_BYTES = 256;

template <typename K>
__device__ __forceinline__ void better(K& k, int& r, K k2, int r2) {
  if (k2 > k || (k2 == k && r2 < r)) {
    k = k2;
    r = r2;
  }
}

__device__ __forceinline__ void put(u64* p, unsigned payload, unsigned tag) {
  __hip_atomic_store(p, (static_cast<u64>(payload) << 32) | tag, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 get(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
// Reads the published row at `src` (n elements, 2 words each) into LDS `dst`, and, when `dst2`
// is set, the row at `src2` into `dst2`.  Lane l takes elements l, l + 64, ...; every load of
// a lane is issued before any tag is checked, so a row that is already visible costs one round
// trip, and a round with a stale word re-reads the whole set.
// The waits of a wave count their poll rounds together, over the whole launch; every 32nd
// round reads the abort pair.  A wave that is past the limit stores `id` (the launch's first
// tag, never 0, so a cleared word names no launch) in word 0 and its wait returns false; so
// does a wait that finds `id` there (another workgroup of the launch gave up).  The other
// rounds only count, so a poll loop keeps no more than the counter in registers.
__device__ __forceinline__ bool wait_expired(const u64* keys, unsigned id, unsigned rounds) {
  u64* pair = const_cast<u64*>(keys) + PANEL_ABORT_OFF / sizeof(u64);
  const u64 w = get(pair);
  if (static_cast<unsigned>(w) == id) return true;
  const unsigned lim = static_cast<unsigned>(w >> 32);
  if (rounds < (lim ? lim : DEFAULT_WAIT_LIMIT)) return false;
  __hip_atomic_store(reinterpret_cast<unsigned*>(pair), id, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  return true;
}
__device__ __forceinline__ bool keep_waiting(const u64* keys, unsigned id, unsigned& rounds) {
  ++rounds;
  return (rounds & 31u) != 0 || !wait_expired(keys, id, rounds);
}

template <typename T, int CPL>
__device__ __forceinline__ bool fetch_rows(const u64* src, const u64* src2, int n, int lane,
                                           unsigned tag, T* dst, T* dst2, const u64* area,
                                           unsigned id, unsigned& rounds) {
  constexpr int EW = sizeof(T) == 8 ? 2 : 1;
  u64 a[CPL][EW], b[CPL][EW];
  const bool two = src2 != nullptr;
  for (;;) {
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      const int c = lane + 64 * j;
      if (c < n) {
#pragma unroll
        for (int e = 0; e < EW; ++e) a[j][e] = get(src + 2 * c + e);
        if (two) {
#pragma unroll
          for (int e = 0; e < EW; ++e) b[j][e] = get(src2 + 2 * c + e);
        }
      }
    }
    bool ok = true;
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      if (lane + 64 * j < n) {
#pragma unroll
        for (int e = 0; e < EW; ++e) {
          ok = ok && static_cast<unsigned>(a[j][e]) == tag;
          if (two) ok = ok && static_cast<unsigned>(b[j][e]) == tag;
        }
      }
    }
    if (__all(ok)) break;  // wave-uniform loop: the round counter stays scalar
    if (!keep_waiting(area, id, rounds)) return false;
    __builtin_amdgcn_s_sleep(1);
  }
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    const int c = lane + 64 * j;
    if (c < n) {
      if constexpr (sizeof(T) == 8) {
        dst[c] = __builtin_bit_cast(T, ((a[j][0] >> 32) << 32) | (a[j][1] >> 32));
        if (two) dst2[c] = __builtin_bit_cast(T, ((b[j][0] >> 32) << 32) | (b[j][1] >> 32));
      } else {
        dst[c] = __builtin_bit_cast(T, static_cast<unsigned>(a[j][0] >> 32));
        if (two) dst2[c] = __builtin_bit_cast(T, static_cast<unsigned>(b[j][0] >> 32));
      }
    }
  }
  return true;
}

// Polls the W key slots (slot q at keys + q KSTRIDE: the key in one tagged word for FP32, two
// for FP64, then the packed label) until all carry `tag`, and reduces them per lane into
// (k, r).  Lane l reads slots l, l + 64, ... (S per lane); all reads of a round are issued
// before any tag is checked, and they are unconditional (a lane's unused slots read slot
// W - 1 and are not checked): reads under a per-slot branch would each be merged into fresh
// registers and waited for one group at a time.
template <typename K, int S>
__device__ __forceinline__ bool poll_keys(const u64* keys, int W, int lane, unsigned tag, K& k,
                                          int& r, const u64* area, unsigned id,
                                          unsigned& rounds) {
  constexpr int EW = sizeof(K) == 8 ? 2 : 1;  // words of the key
  u64 a[S], b[S], c[S];
  for (;;) {
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const u64* kp = keys + (size_t)min(lane + 64 * j, W - 1) * KSTRIDE;
      a[j] = get(kp);
      if constexpr (EW == 2) b[j] = get(kp + 1);
      c[j] = get(kp + EW);
    }
    bool ok = true;
#pragma unroll
    for (int j = 0; j < S; ++j)
      ok = ok && (lane + 64 * j >= W ||
                  (static_cast<unsigned>(a[j]) == tag && static_cast<unsigned>(c[j]) == tag &&
                   (EW == 1 || static_cast<unsigned>(b[j]) == tag)));
    if (__all(ok)) break;  // wave-uniform loop: the round counter stays scalar
    if (!keep_waiting(area, id, rounds)) return false;
    __builtin_amdgcn_s_sleep(1);
  }
  k = K(-2);
  r = INT_MAX;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    if (lane + 64 * j < W) {
      K kj;
      if constexpr (EW == 2)
        kj = __builtin_bit_cast(K, ((a[j] >> 32) << 32) | (b[j] >> 32));
      else
        kj = __builtin_bit_cast(K, static_cast<unsigned>(a[j] >> 32));
      better(k, r, kj, static_cast<int>(c[j] >> 32));
    }
  }
  return true;
}

// Selects element k of a register array with a uniform run-time index (unrolled compares,
// no dynamic register indexing).  The empty asm hides each element's origin: otherwise
// the select chain over array loads folds into one load at a variable address, and the
// whole array is demoted to scratch memory.
template <typename X>
__device__ __forceinline__ X opaque(X x) {
  asm volatile("" : "+v"(x));
  return x;
}
template <typename X, int N>
__device__ __forceinline__ X reg_get(const X (&v)[N], int k) {
  X x = opaque(v[0]);
#pragma unroll
  for (int j = 1; j < N; ++j) x = (j == k) ? opaque(v[j]) : x;
  return x;
}
template <typename X, int N>
__device__ __forceinline__ void reg_set(X (&v)[N], int k, X x) {
#pragma unroll
  for (int j = 0; j < N; ++j) v[j] = opaque((j == k) ? x : v[j]);
}

// The workgroup's R panel rows live in registers: local row lr = wave + NW k (k < R/NW) is
// held by the 64 lanes of wave lr % NW, lane l holding columns l and l + 64; the row's label
// sits in LDS (lane k of the wave reads row k's for the arg-max) and, the same in every lane, in
// registers.  Column i of a wave's rows therefore sits in one lane.
//
// Column c of a wave's rows (lane c & 63 of v0 or v1) goes to colb[0, RW) in LDS: the arg-max
// over the rows is then a lane-parallel reduction instead of a scan in one lane, and the rank-1
// update reads the rows' multipliers from there as broadcasts.  Rows whose label is <= thr
// (already pivotal) are stored as zero, so that they take a zero multiplier.  (Rows below the
// panel's last row hold zeros from the start; the arg-max and the write-back skip them.)
template <typename T, int RW>
__device__ __forceinline__ void column_to_lds(const T (&v0)[RW], const T (&v1)[RW], int c,
                                              int lane, const int (&lbl)[RW], int thr,
                                              T* colb) {
  if (lane == (c & 63)) {
    // 16-byte stores (colb is 16-byte aligned, RW a multiple of the vector width)
    constexpr int VN = 16 / sizeof(T);
    typedef T V __attribute__((ext_vector_type(VN)));
#pragma unroll
    for (int k0 = 0; k0 < RW; k0 += VN) {
      V y;
#pragma unroll
      for (int e = 0; e < VN; ++e) {
        const T x = c < 64 ? v0[k0 + e] : v1[k0 + e];
        y[e] = lbl[k0 + e] > thr ? x : T(0);
      }
      *reinterpret_cast<V*>(colb + k0) = y;
    }
  }
}

// Arg-max of |colb[k]| over the wave's candidate rows (k < kmax, label >= i): the largest
// value wins and the smallest label breaks ties (the first index in the row order of the
// reference, whose rows are physically interchanged), a NaN entry counts as -1 (below every
// number); key -2 and INT_MAX when the wave has no candidate.  Labels are packed as
// label * 256 + workgroup.  Every lane ends with the key, the packed label, the slot k of the
// winner and the slot of the row labelled i (-1 when the wave does not hold it).
template <typename T, int RW>
__device__ __forceinline__ void wave_argmax(const T* colb, const int* lblw, int i, int kmax, int w,
                                            int lane, T& key, int& pk, int& kbest,
                                            int& kdiag) {
  // the column was written by one lane of this wave: wait for the LDS writes (in-order
  // LDS queue per wave), and keep the compiler from hoisting the reads above them
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  const T x = lane < RW ? colb[lane] : T(0);
  const int L = lane < RW ? lblw[lane] : INT_MAX;
  const bool cand = lane < kmax && L >= i;
  T a = cand ? fmax(fabs(x), T(-1)) : T(-2);
  const int mine = cand ? L * 256 + w : INT_MAX;
  int k = mine;
  lanes_argmax_dpp<RW>(a, k);  // the rows sit in lanes [0, RW)
  key = a;
  pk = k;
  const unsigned long long bw = __ballot(cand && mine == k);
  kbest = bw ? __ffsll(bw) - 1 : -1;
  const unsigned long long bd = __ballot(lane < kmax && L == i);
  kdiag = bd ? __ffsll(bd) - 1 : -1;
}

// Step i of the rank-1 update on one wave's rows: the multiplier of a row is its column-i
// entry times 1/pivot, stored in place; columns > i take fma(-l, u, a) (GeR), columns < i keep
// their values (u is zero there, so the fma returns them unchanged).  The column-i entries
// come from colb, where the rows that are already pivotal and the pivot row of this step hold
// zero, so every lane forms the same multipliers from LDS broadcasts and those rows take
// multiplier 0 without a per-row test; only the in-place store in the lane of column i is
// limited to the rows labelled > i, by one compare per row against a lane value.  Column i + 1
// of the rows then goes to LDS for the next arg-max.  HI: column i is in the upper half (v1).
template <typename T, int RW, bool HI>
__device__ __forceinline__ void panel_update(T (&v0)[RW], T (&v1)[RW], T* colb,
                                             const int (&lbl)[RW], int i, T rinv, T u0, T u1,
                                             int lane) {
  const int cl = HI ? i - 64 : i;  // lane of column i
  const int q = lane == cl ? i : INT_MAX;
  T l[RW];
#pragma unroll
  for (int k = 0; k < RW; ++k) l[k] = colb[k] * rinv;
  if constexpr (!HI) {
    const T z0 = lane > i ? u0 : T(0);
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const T n0 = fma(-l[k], z0, v0[k]);
      v1[k] = fma(-l[k], u1, v1[k]);  // every column >= 64 > i
      v0[k] = lbl[k] > q ? l[k] : n0;
    }
  } else {
    const T z1 = lane > cl ? u1 : T(0);
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const T n1 = fma(-l[k], z1, v1[k]);
      v1[k] = lbl[k] > q ? l[k] : n1;
    }
  }
  // rows labelled <= i are pivotal from now on
  column_to_lds(v0, v1, i + 1, lane, lbl, i, colb);
}

template <typename T, int R, int NW>
__global__ __launch_bounds__(NW * 64) void getf2_panel_kernel(int m, int nb, T* __restrict__ P,
                                                          int64_t ld, int32_t* __restrict__ piv,
                                                          int32_t* __restrict__ info,
                                                          PanelSync sy, unsigned base,
                                                          int info_base) {
  constexpr int RW = R / NW;  // rows per wave
  __shared__ T prow[NB_MAX];
  __shared__ __attribute__((aligned(16))) T colb[NW][RW];    // per wave: the next column
  __shared__ __attribute__((aligned(16))) int lblw[NW][RW];  // per wave: labels of its rows
  __shared__ T sk[NW];  // per wave: local arg-max key, packed label, slot
  __shared__ int sr[NW], ss[NW];
  __shared__ int swin;

  const int W = gridDim.x;
  const int w = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r0 = w * R;

  if (w == 0 && tid == 0) {
    // ordered before step 0's publication, which every workgroup waits for before its
    // final atomicAdd on the counter
    __hip_atomic_store(sy.list_count, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __atomic_thread_fence(__ATOMIC_RELEASE);
  }
  // an earlier panel already hit a zero pivot: leave everything as it is (empty list)
  if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;

  T v0[RW], v1[RW];
  const int c0 = lane, c1 = lane + 64;
  const int gr0 = r0 + wave;
#pragma unroll
  for (int k = 0; k < RW; ++k) {
    const int gr = gr0 + NW * k;
    const bool ok = gr < m;
    v0[k] = (ok && c0 < nb) ? P[(int64_t)gr * ld + c0] : T(0);
    v1[k] = (ok && c1 < nb) ? P[(int64_t)gr * ld + c1] : T(0);
  }
  const int kmax = min(max((m + NW - 1 - gr0) / NW, 0), RW);  // rows k < kmax lie below m
  // the wave's row labels: in LDS for the arg-max (lane k reads row k's) and, the same in every
  // lane, in registers for the update
  int lbl[RW];
#pragma unroll
  for (int k = 0; k < RW; ++k) lbl[k] = gr0 + NW * k;
  if (lane < RW) lblw[wave][lane] = gr0 + NW * lane;
  column_to_lds(v0, v1, 0, lane, lbl, -1, colb[wave]);

  unsigned rounds = 0;  // wave 0's poll rounds in this launch (every wait counts)
  for (int i = 0; i < nb; ++i) {
    const int par = i & 1;
    const unsigned tag = base + static_cast<unsigned>(i) + 1u;
    // 1. local arg-max of column i over the rows not yet pivotal: per wave, then across waves
    T key;
    int pk, kd;
    {
      int kb;
      wave_argmax<T, RW>(colb[wave], lblw[wave], i, kmax, w, lane, key, pk, kb, kd);
      if (lane == 0) {
        sk[wave] = key;
        sr[wave] = pk;
        ss[wave] = kb;
      }
    }
    kd = __builtin_amdgcn_readfirstlane(kd);
    __syncthreads();
    // the waves' candidates in lanes [0, NW), reduced by DPP; qb = the wave that holds the
    // winner (its packed label is unique), kc = its slot there
    int qb;
    {
      const int pq = lane < NW ? sr[lane] : INT_MAX;
      key = lane < NW ? sk[lane] : T(-2);
      pk = pq;
      lanes_argmax_dpp<NW>(key, pk);
      pk = __builtin_amdgcn_readfirstlane(pk);
      const unsigned long long bq = __ballot(lane < NW && pq == pk);
      qb = __builtin_amdgcn_readfirstlane(bq ? __ffsll(bq) - 1 : 0);
    }
    const int kc = __builtin_amdgcn_readfirstlane(ss[qb]);  // slot of the candidate in wave qb
    // 2. publish the candidate row and the key
    if (key > T(-2) && wave == qb) {
      u64* my = sy.cand + ((size_t)par * WMAX + w) * NB_MAX * 2;
      put_elem<T>(my + 2 * c0, reg_get(v0, kc), tag);
      put_elem<T>(my + 2 * c1, reg_get(v1, kc), tag);
    }
    if (tid == 0) {
      u64* kp = sy.keys + ((size_t)par * WMAX + w) * KSTRIDE;
      if constexpr (sizeof(T) == 8) {
        const u64 kb = __builtin_bit_cast(u64, key);
        put(kp, static_cast<unsigned>(kb >> 32), tag);
        put(kp + 1, static_cast<unsigned>(kb), tag);
        put(kp + 2, static_cast<unsigned>(pk), tag);
      } else {
        put(kp, __builtin_bit_cast(unsigned, key), tag);
        put(kp + 1, static_cast<unsigned>(pk), tag);
      }
    }
    // 3. global winner (identical in every workgroup): wave 0 reads the W slots, lane q the
    //    slots q, q + 64, ...; all requests of a lane are in flight together
    if (wave == 0) {
      T gk;
      int gp;
      const u64* ks = sy.keys + (size_t)par * WMAX * KSTRIDE;
      const int nj = (W + 63) >> 6;
      bool ok;
      if (nj <= 1)
        ok = poll_keys<T, 1>(ks, W, lane, tag, gk, gp, sy.keys, base + 1u, rounds);
      else if (nj == 2)
        ok = poll_keys<T, 2>(ks, W, lane, tag, gk, gp, sy.keys, base + 1u, rounds);
      else if (nj == 3)
        ok = poll_keys<T, 3>(ks, W, lane, tag, gk, gp, sy.keys, base + 1u, rounds);
      else
        ok = poll_keys<T, 4>(ks, W, lane, tag, gk, gp, sy.keys, base + 1u, rounds);
      if (ok) {
        wave_argmax_dpp(gk, gp);  // DPP across the lanes (every lane ends with the winner)
        gp = __builtin_amdgcn_readfirstlane(gp);
        // the winner's row, read by this wave right away, to LDS for the others
        const u64* cw = sy.cand + ((size_t)par * WMAX + (gp & 255)) * NB_MAX * 2;
        ok = fetch_rows<T, NB_MAX / 64>(cw, nullptr, NB_MAX, lane, tag, prow, prow, sy.keys,
                                        base + 1u, rounds);
      }
      if (lane == 0) swin = ok ? gp : -1;
    }
    __syncthreads();  // prow and the winner in LDS
    const int gp = __builtin_amdgcn_readfirstlane(swin);
    // a wait gave up: leave the column loop as on a zero pivot (the rows are written back in
    // their current, consistent order; the host reports the failure)
    if (gp < 0) break;
    const int p = gp >> 8;  // label of the pivot row
    const T pv = prow[i];
    if (w == 0 && tid == 0) piv[i] = p - i;
    const bool stop = (pv == static_cast<T>(0));
    if (stop && w == 0 && tid == 0) atomicCAS(info, 0, info_base + i + 1);
    // 4. interchange rows i and p by their labels (no data moves): the pivot row takes label
    //    i and, from now on, no multiplier; the row labelled i takes label p
    if (w == (gp & 255) && wave == qb) {
      reg_set(lbl, kc, i);
      if (lane == 0) {
        lblw[wave][kc] = i;
        colb[wave][kc] = T(0);
      }
    }
    if (kd >= 0 && p != i) {
      reg_set(lbl, kd, p);
      if (lane == 0) lblw[wave][kd] = p;
    }
    if (stop) break;
    // 5. scale column i below the pivot, update the rest of every row and track the
    //    arg-max of column i + 1 over the updated rows
    const T rinv = static_cast<T>(1) / pv;
    const T u0 = prow[c0], u1 = prow[c1];
    if (i < 64) {
      panel_update<T, RW, false>(v0, v1, colb[wave], lbl, i, rinv, u0, u1, lane);
    } else {
      panel_update<T, RW, true>(v0, v1, colb[wave], lbl, i, rinv, u0, u1, lane);
    }
  }

  // write every row to the position of its label and emit the row movement list
  // (dst <- src, panel-relative) for the rest of the matrix
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): this wave's label writes
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int k = 0; k < RW; ++k) {
    if (k < kmax) {
      const int dst = lblw[wave][k];
      if (c0 < nb) P[(int64_t)dst * ld + c0] = v0[k];
      if (c1 < nb) P[(int64_t)dst * ld + c1] = v1[k];
    }
  }
  if (lane < kmax) {
    const int src = gr0 + NW * lane, dst = lblw[wave][lane];
    if (dst != src) {
      const int slot = atomicAdd(sy.list_count, 1);
      sy.list[2 * slot] = dst;
      sy.list[2 * slot + 1] = src;
    }
  }
}

// ---------------------------------------------------------------------------------------
// FP32 panel with one row per lane (getf2_rows_kernel): the same steps, pivots, labels and
// outputs as getf2_panel_kernel<float>, another distribution of the panel.  Lane l of wave v
// of workgroup w holds panel row w * 64 NW + 64 v + l, all of its (at most 128) columns, in
// registers: the arg-max of a column is a reduction across lanes, and the rank-1 update of a
// row is a run of fused multiply-adds against the pivot row broadcast from LDS.  With 64 NW
// rows per workgroup few workgroups hold the panel (32 at m = 16384), and each publishes its
// candidate row together with its key, so a column step costs one round trip between them.
// Column step i:
//   1. arg-max of column i: an unsigned key per lane (0: not a candidate, 1: NaN, else the
//      bits of |x| + 2, so NaN ranks below every number as in AMax) reduced by DPP max across
//      the lanes, ties broken by the smallest label (a DPP min only when several lanes tie);
//      the waves' winners meet in an LDS atomic max on (key << 32 | ~label), one barrier;
//   2. the wave that holds the workgroup's candidate stages that row in LDS and publishes key,
//      packed label (label * 256 + workgroup) and the row's columns from the current 16-column
//      block on as tagged words in the workgroup's slot (no candidate: key 0 and zeros);
//   3. wave v reads slots v, v + NW, ... -- keys in lanes [0, S), rows in every lane, all
//      loads in flight before any tag is checked -- keeps the best, writes its row to LDS
//      (columns <= i zeroed, the column-i value aside) and its key to a second LDS atomic max;
//   4. after a barrier every workgroup reads the same winner; rows are relabelled (no data
//      moves): the pivot row takes label i, the row labelled i the pivot's label;
//   5. every row not yet pivotal forms its multiplier (column i times 1/pivot) and updates
//      column i + 1 only -- the next arg-max needs no more.  The update of its other columns
//      after i + 1 is applied in step i + 1, between issuing the slot reads and checking them,
//      so it runs in the shadow of the exchange; the candidate row published in step i + 1
//      takes that pending update while it is staged (the same fma on the same operands).
// The update of a column by a pivot row is fma(-l, u, a) in the order of the pivots, as in
// the reference; columns before the pivot keep their values (a select, not an fma with zero).
// Columns are addressed by the static block (one code copy per 16-column block) and by the
// uniform step inside the block, a dynamic element index of a 16-wide vector.  With a single
// workgroup (m <= 64 NW) nothing is published: the candidate row goes to LDS directly.
typedef float f16v __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));

// slot of a workgroup: word 0 the key, word 1 the packed label, words 8 .. 135 the row's 128
// values; 144 words keep slots on 128-byte lines
constexpr int RS_WORDS = 144;
static_assert(2 * (size_t)WMAX * RS_WORDS * sizeof(u64) <= PANEL_CAND_BYTES,
              "row-panel slots exceed the candidate area");

template <int NW, int S, bool TWO>
struct RowsPanel {
  // uniform over the workgroup
  int m, nb, W, w, wave, lane, row;
  unsigned base, rounds;
  int qwp;  // wrow slot of the previous step's pivot row
  int32_t* piv;
  int32_t* info;
  int info_base;
  PanelSync sy;
  float (*stage)[NB_MAX + 4];      // per wave: the raw candidate row, its pending multiplier
  float (*wrow)[NW][NB_MAX];       // [parity][wave]: the best row the wave read (cols <= i: 0)
  float (*pval)[NW];               // [parity][wave]: its column-i value
  unsigned long long* best;        // [parity]: workgroup winner, [2 + parity]: global winner
  int* gaveup;
  // per lane: the row, its label (INT_MAX beyond m), its column-i value and its multiplier of
  // the previous step (the update of its columns after i + 1 still pending)
  f16v blk[NB_MAX / 16];
  int lbl;
  float cur, lprev;

  // the previous step's update of the columns after i (rows that took part in it)
  template <int JB>
  __device__ __forceinline__ void apply_prev(int i) {
    if (i == 0 || !(lbl >= i && lbl != INT_MAX)) return;
    const float nl = -lprev;
    const float* u = wrow[(i - 1) & 1][qwp];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const f4v uu = *reinterpret_cast<const f4v*>(u + 16 * JB + 4 * q);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float y = fmaf(nl, uu[e], blk[JB][4 * q + e]);
        blk[JB][4 * q + e] = 16 * JB + 4 * q + e > i ? y : blk[JB][4 * q + e];
      }
    }
    // one 16-column block at a time: the scheduler would otherwise hoist every LDS read of
    // the pivot row ahead of the multiply-adds and run out of registers
#pragma unroll
    for (int B = JB + 1; B < NB_MAX / 16; ++B) {
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const f4v uu = *reinterpret_cast<const f4v*>(u + 16 * B + 4 * q);
#pragma unroll
        for (int e = 0; e < 4; ++e) blk[B][4 * q + e] = fmaf(nl, uu[e], blk[B][4 * q + e]);
      }
    }
  }

  // The candidate row (label lab, held by this wave) with its pending update applied: lane l
  // ends with columns 2l and 2l + 1 (lanes >= 8 JB; zeros when has is false).
  template <int JB>
  __device__ __forceinline__ float2 staged_pair(int i, bool has, int lab) {
    if (has && lbl == lab) {
#pragma unroll
      for (int B = JB; B < NB_MAX / 16; ++B)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          f4v x;
#pragma unroll
          for (int e = 0; e < 4; ++e) x[e] = blk[B][4 * q + e];
          *reinterpret_cast<f4v*>(&stage[wave][16 * B + 4 * q]) = x;
        }
      stage[wave][NB_MAX] = lprev;
    }
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): the staged row
    __builtin_amdgcn_wave_barrier();
    float2 x = make_float2(0.f, 0.f);
    if (has && lane >= 8 * JB) {
      const int c0 = 2 * lane;
      x = *reinterpret_cast<const float2*>(&stage[wave][c0]);
      if (i > 0) {
        const float nl = -stage[wave][NB_MAX];
        const float* u = wrow[(i - 1) & 1][qwp];
        const float2 uu = *reinterpret_cast<const float2*>(u + c0);
        if (c0 > i) x.x = fmaf(nl, uu.x, x.x);
        if (c0 + 1 > i) x.y = fmaf(nl, uu.y, x.y);
      }
    }
    return x;
  }

  // One column step; true when the panel stops (zero pivot, or a wait gave up).
  template <int JB>
  __device__ __forceinline__ bool step(int ii) {
    const int i = 16 * JB + ii;
    const int par = i & 1;
    const unsigned tag = base + static_cast<unsigned>(i) + 1u;
    // 1. the workgroup's candidate
    {
      const bool cand = lbl >= i && lbl != INT_MAX;
      const unsigned key = arg_key(cur, cand);
      unsigned kw;
      const int wl = lanes_winner<64>(key, static_cast<unsigned>(lbl), lane, kw);
      if (wl >= 0 && lane == 0) {
        const unsigned lab = __builtin_amdgcn_readlane(lbl, wl);
        __hip_atomic_fetch_max(&best[par], (static_cast<u64>(kw) << 32) | (0xFFFFFFFFu - lab),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    __syncthreads();
    const u64 wb = best[par];
    if (wave == 0 && lane == 0) {
      // the words of the other parity were last read before this barrier
      best[par ^ 1] = 0;
      best[2 + (par ^ 1)] = 0;
    }
    const unsigned kwg = static_cast<unsigned>(wb >> 32);
    const bool has = kwg != 0u;
    const int labwg = has ? static_cast<int>(0xFFFFFFFFu - static_cast<unsigned>(wb)) : INT_MAX;
    const bool holder = __ballot(has && lbl == labwg) != 0;  // this wave holds the candidate
    bool gave = false;
    // 2. publish (several workgroups) or stage (one): the candidate row with its pending
    //    update, from the wave that holds it
    if constexpr (TWO) {
      // 3. (two-phase exchange) wave 0 reads every slot's key (lane j: slot j), then the
      //    winner's row into wrow slot 0
      const bool poller = W > 1 && wave == 0;
      const bool rl = lane >= 8 * JB;        // lanes that publish columns from block JB on
      const bool rf = 2 * lane + 1 >= i;     // lanes that fetch column i or a later one
      const u64* area = sy.cand + (size_t)par * WMAX * RS_WORDS;
      u64 kw0 = 0, kw1 = 0;
      auto issue = [&]() {
        const u64* ks = area + (size_t)min(lane, W - 1) * RS_WORDS;
        kw0 = get(ks);
        kw1 = get(ks + 1);
      };
      if (W > 1) {
        if (holder || (!has && wave == 0)) {
          const float2 x = staged_pair<JB>(i, has, labwg);
          u64* slot = sy.cand + ((size_t)par * WMAX + w) * RS_WORDS;
          if (rl) {
            put(slot + 8 + 2 * lane, __builtin_bit_cast(unsigned, x.x), tag);
            put(slot + 9 + 2 * lane, __builtin_bit_cast(unsigned, x.y), tag);
          }
          if (lane == 0) {
            put(slot, kwg, tag);
            put(slot + 1, static_cast<unsigned>(has ? labwg * 256 + w : INT_MAX), tag);
          }
        }
        if (poller) issue();
      } else {
        if (holder) {
          const float2 x = staged_pair<JB>(i, true, labwg);
          if (rl) {
            const int c0 = 2 * lane;
            *reinterpret_cast<float2*>(&wrow[par][0][c0]) =
                make_float2(c0 > i ? x.x : 0.f, c0 + 1 > i ? x.y : 0.f);
            if (c0 == i) pval[par][0] = x.x;
            if (c0 + 1 == i) pval[par][0] = x.y;
          }
        }
        if (has && wave == 0 && lane == 0)
          best[2 + par] = (static_cast<u64>(kwg) << 32) |
                          (0xFFFFFFFFu - static_cast<unsigned>(labwg) * 256u);
      }
      apply_prev<JB>(i);
      if (poller) {
        // While keys are missing, the row of the best key seen so far is read with the next
        // round of keys: when that candidate wins, its row is already in registers.
        int qs = -1;  // slot whose row a0 / a1 hold (read in the same round as the keys)
        u64 a0 = 0, a1 = 0;
        auto row_of = [&](int q) { return rf ? area + (size_t)q * RS_WORDS + 8 + 2 * lane
                                             : area + (size_t)q * RS_WORDS; };
        for (;;) {
          const bool have = lane >= W || (static_cast<unsigned>(kw0) == tag &&
                                          static_cast<unsigned>(kw1) == tag);
          if (__all(have)) break;
          if (!keep_waiting(sy.keys, base + 1u, rounds)) {
            gave = true;
            break;
          }
          unsigned kq;
          const unsigned kj = lane < W && have ? static_cast<unsigned>(kw0 >> 32) : 0u;
          const unsigned pj = lane < W && have ? static_cast<unsigned>(kw1 >> 32) : 0xFFFFFFFFu;
          const int jq = lanes_winner<64>(kj, pj, lane, kq);
          qs = jq >= 0 ? static_cast<int>(__builtin_amdgcn_readlane(pj, jq) & 255u) : -1;
          issue();
          if (qs >= 0) {
            const u64* vs = row_of(qs);
            a0 = get(vs);
            a1 = get(vs + 1);
          }
        }
        if (!gave) {
          unsigned kb;
          const unsigned kj = lane < W ? static_cast<unsigned>(kw0 >> 32) : 0u;
          const unsigned pj = lane < W ? static_cast<unsigned>(kw1 >> 32) : 0xFFFFFFFFu;
          const int jb = lanes_winner<64>(kj, pj, lane, kb);
          if (jb >= 0) {
            const unsigned pk = __builtin_amdgcn_readlane(pj, jb);
            const int q = static_cast<int>(pk & 255u);
            const u64* vs = row_of(q);
            if (q != qs) {
              a0 = get(vs);
              a1 = get(vs + 1);
            }
            while (!__all(!rf || (static_cast<unsigned>(a0) == tag &&
                                  static_cast<unsigned>(a1) == tag))) {
              if (!keep_waiting(sy.keys, base + 1u, rounds)) {
                gave = true;
                break;
              }
              __builtin_amdgcn_s_sleep(1);
              a0 = get(vs);
              a1 = get(vs + 1);
            }
            if (!gave) {
              if (rf) {
                const float x0 = __builtin_bit_cast(float, static_cast<unsigned>(a0 >> 32));
                const float x1 = __builtin_bit_cast(float, static_cast<unsigned>(a1 >> 32));
                const int c0 = 2 * lane;
                *reinterpret_cast<float2*>(&wrow[par][0][c0]) =
                    make_float2(c0 > i ? x0 : 0.f, c0 + 1 > i ? x1 : 0.f);
                if (c0 == i) pval[par][0] = x0;
                if (c0 + 1 == i) pval[par][0] = x1;
              }
              if (lane == 0) best[2 + par] = (static_cast<u64>(kb) << 32) | (0xFFFFFFFFu - pk);
            }
          }
        }
        if (gave && lane == 0) *gaveup = 1;
      }
    } else {
      const int nv = W > 1 && wave < W ? (W - wave + NW - 1) / NW : 0;  // slots this wave reads
      const bool rl = lane >= 8 * JB;  // lanes that hold columns from block JB on
      const u64* area = sy.cand + (size_t)par * WMAX * RS_WORDS;
      u64 kw0 = 0, kw1 = 0, va[S], vb[S];
      // every lane reads its pair of every slot (lanes < 8 JB too: their words are not checked)
      auto issue = [&]() {
        const u64* ks = area + (size_t)(wave + NW * min(lane, nv - 1)) * RS_WORDS;
        kw0 = get(ks);
        kw1 = get(ks + 1);
  #pragma unroll
        for (int j = 0; j < S; ++j) {
          const u64* vs = area + (size_t)(wave + NW * min(j, nv - 1)) * RS_WORDS + 8 + 2 * lane;
          va[j] = get(vs);
          vb[j] = get(vs + 1);
        }
      };
      if (W > 1) {
        if (holder || (!has && wave == 0)) {
          const float2 x = staged_pair<JB>(i, has, labwg);
          u64* slot = sy.cand + ((size_t)par * WMAX + w) * RS_WORDS;
          if (rl) {
            put(slot + 8 + 2 * lane, __builtin_bit_cast(unsigned, x.x), tag);
            put(slot + 9 + 2 * lane, __builtin_bit_cast(unsigned, x.y), tag);
          }
          if (lane == 0) {
            put(slot, kwg, tag);
            put(slot + 1, static_cast<unsigned>(has ? labwg * 256 + w : INT_MAX), tag);
          }
        }
        // 3. the first reads of this wave's slots go out before the pending update
        if (nv > 0) issue();
      } else {
        if (holder) {
          const float2 x = staged_pair<JB>(i, true, labwg);
          if (rl) {
            const int c0 = 2 * lane;
            *reinterpret_cast<float2*>(&wrow[par][0][c0]) =
                make_float2(c0 > i ? x.x : 0.f, c0 + 1 > i ? x.y : 0.f);
            if (c0 == i) pval[par][0] = x.x;
            if (c0 + 1 == i) pval[par][0] = x.y;
          }
        }
        if (has && wave == 0 && lane == 0)
          best[2 + par] = (static_cast<u64>(kwg) << 32) |
                          (0xFFFFFFFFu - static_cast<unsigned>(labwg) * 256u);
      }
      apply_prev<JB>(i);
      if (nv > 0) {
        for (;;) {
          bool ok = lane >= nv || (static_cast<unsigned>(kw0) == tag &&
                                   static_cast<unsigned>(kw1) == tag);
  #pragma unroll
          for (int j = 0; j < S; ++j)
            ok = ok && (!rl || j >= nv ||
                        (static_cast<unsigned>(va[j]) == tag && static_cast<unsigned>(vb[j]) == tag));
          if (__all(ok)) break;  // wave-uniform loop: the round counter stays scalar
          if (!keep_waiting(sy.keys, base + 1u, rounds)) {
            gave = true;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          issue();
        }
        if (!gave) {
          unsigned kb;
          const unsigned kj = lane < nv ? static_cast<unsigned>(kw0 >> 32) : 0u;
          const unsigned pj = lane < nv ? static_cast<unsigned>(kw1 >> 32) : 0xFFFFFFFFu;
          const int jb = lanes_winner<S>(kj, pj, lane, kb);
          if (jb >= 0) {
            const unsigned pk = __builtin_amdgcn_readlane(pj, jb);
            if (rl) {
              const float x0 = __builtin_bit_cast(float, static_cast<unsigned>(reg_get(va, jb) >> 32));
              const float x1 = __builtin_bit_cast(float, static_cast<unsigned>(reg_get(vb, jb) >> 32));
              const int c0 = 2 * lane;
              *reinterpret_cast<float2*>(&wrow[par][wave][c0]) =
                  make_float2(c0 > i ? x0 : 0.f, c0 + 1 > i ? x1 : 0.f);
              if (c0 == i) pval[par][wave] = x0;
              if (c0 + 1 == i) pval[par][wave] = x1;
            }
            if (lane == 0)
              __hip_atomic_fetch_max(&best[2 + par], (static_cast<u64>(kb) << 32) | (0xFFFFFFFFu - pk),
                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        } else if (lane == 0) {
          *gaveup = 1;
        }
      }
    }
    __syncthreads();
    // 4. the winner, the same in every workgroup
    if (*gaveup) return true;
    const u64 gb = best[2 + par];
    if (static_cast<unsigned>(gb >> 32) == 0u) return true;  // no candidate (m >= nb: never)
    const unsigned gp = 0xFFFFFFFFu - static_cast<unsigned>(gb);
    const int p = static_cast<int>(gp >> 8);  // label of the pivot row
    // wrow slot of the pivot row: the wave that read the owner's slot, or wave 0 (two-phase
    // exchange, and one workgroup, whose owner is 0)
    const int qw = TWO ? 0 : static_cast<int>(gp & 255u) % NW;
    const float pv = pval[par][qw];
    if (w == 0 && wave == 0 && lane == 0) piv[i] = p - i;
    const bool stop = pv == 0.f;
    if (stop && w == 0 && wave == 0 && lane == 0) atomicCAS(info, 0, info_base + i + 1);
    lbl = lbl == i ? p : (lbl == p ? i : lbl);
    qwp = qw;
    if (stop) return true;
    // 5. multiplier, and the update of column i + 1 for the next arg-max
    constexpr int JN = JB + 1 < NB_MAX / 16 ? JB + 1 : JB;
    if (lbl > i && lbl != INT_MAX) {
      const float l = cur * (1.f / pv);
      blk[JB][ii] = l;
      lprev = l;
      if (i + 1 < nb) {
        const float un = wrow[par][qw][i + 1];
        if (ii < 15) {
          const float y = fmaf(-l, un, blk[JB][min(ii + 1, 15)]);
          blk[JB][min(ii + 1, 15)] = y;
          cur = y;
        } else {
          blk[JN][0] = fmaf(-l, un, blk[JN][0]);
          cur = blk[JN][0];
        }
      }
    } else {
      cur = ii < 15 ? blk[JB][min(ii + 1, 15)] : blk[JN][0];
    }
    return false;
  }

  template <int JB>
  __device__ __forceinline__ bool block() {
    const int ie = min(16, nb - 16 * JB);
    for (int ii = 0; ii < ie; ++ii)
      if (step<JB>(ii)) return true;
    return false;
  }
};

template <int NW, int S, bool TWO>
__global__ __launch_bounds__(NW * 64, 1) void getf2_rows_kernel(int m, int nb,
                                                                float* __restrict__ P, int64_t ld,
                                                                int32_t* __restrict__ piv,
                                                                int32_t* __restrict__ info,
                                                                PanelSync sy, unsigned base,
                                                                int info_base) {
  __shared__ __attribute__((aligned(16))) float stage[NW][NB_MAX + 4];
  __shared__ __attribute__((aligned(16))) float wrow[2][NW][NB_MAX];
  __shared__ float pval[2][NW];
  __shared__ unsigned long long best[4];
  __shared__ int gaveup;

  RowsPanel<NW, S, TWO> st;
  st.m = m;
  st.nb = nb;
  st.W = gridDim.x;
  st.w = blockIdx.x;
  const int tid = threadIdx.x;
  st.lane = tid & 63;
  st.wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  st.row = st.w * 64 * NW + 64 * st.wave + st.lane;
  st.base = base;
  st.rounds = 0;
  st.qwp = 0;
  st.piv = piv;
  st.info = info;
  st.info_base = info_base;
  st.sy = sy;
  st.stage = stage;
  st.wrow = wrow;
  st.pval = pval;
  st.best = best;
  st.gaveup = &gaveup;

  if (st.w == 0 && tid == 0) {
    // ordered before step 0's publication, which every workgroup waits for before its
    // final atomicAdd on the counter
    __hip_atomic_store(sy.list_count, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __atomic_thread_fence(__ATOMIC_RELEASE);
  }
  // an earlier panel already hit a zero pivot: leave everything as it is (empty list)
  if (__builtin_amdgcn_readfirstlane(
          __hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0)
    return;
  if (tid == 0) {
    gaveup = 0;
    for (int k = 0; k < 4; ++k) best[k] = 0;
  }

  const int row = st.row;
  const bool in = row < m;
  st.lbl = in ? row : INT_MAX;
  st.lprev = 0.f;
  float* prow = P + (int64_t)(in ? row : 0) * ld;
  const bool vec = nb == NB_MAX && (ld & 3) == 0 && (reinterpret_cast<uintptr_t>(P) & 15) == 0;
  if (vec) {
#pragma unroll
    for (int q = 0; q < NB_MAX / 4; ++q) {
      const f4v x = in ? reinterpret_cast<const f4v*>(prow)[q] : f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int e = 0; e < 4; ++e) st.blk[q >> 2][(q & 3) * 4 + e] = x[e];
    }
  } else {
#pragma unroll
    for (int c = 0; c < NB_MAX; ++c) st.blk[c >> 4][c & 15] = (in && c < nb) ? prow[c] : 0.f;
  }
  st.cur = st.blk[0][0];
  __syncthreads();  // the LDS words above are cleared before any wave uses them

  // uniform early exits between the blocks: a zero pivot or a wait that gave up
  bool stop = st.template block<0>();
  if (!stop && nb > 16) stop = st.template block<1>();
  if (!stop && nb > 32) stop = st.template block<2>();
  if (!stop && nb > 48) stop = st.template block<3>();
  if (!stop && nb > 64) stop = st.template block<4>();
  if (!stop && nb > 80) stop = st.template block<5>();
  if (!stop && nb > 96) stop = st.template block<6>();
  if (!stop && nb > 112) stop = st.template block<7>();

  // every row to the position of its label, and the row movement list (dst <- src)
  const int lb = st.lbl;
  if (lb != INT_MAX) {
    float* dst = P + (int64_t)lb * ld;
    if (vec) {
#pragma unroll
      for (int q = 0; q < NB_MAX / 4; ++q) {
        f4v x;
#pragma unroll
        for (int e = 0; e < 4; ++e) x[e] = st.blk[q >> 2][(q & 3) * 4 + e];
        reinterpret_cast<f4v*>(dst)[q] = x;
      }
    } else {
#pragma unroll
      for (int c = 0; c < NB_MAX; ++c)
        if (c < nb) dst[c] = st.blk[c >> 4][c & 15];
    }
    if (lb != row) {
      const int slot = atomicAdd(sy.list_count, 1);
      sy.list[2 * slot] = lb;
      sy.list[2 * slot + 1] = row;
    }
  }
}

// ---------------------------------------------------------------------------------------
// FP64 panel with two lanes per row (getf2_rows64_kernel): the design of getf2_rows_kernel for
// FP64, whose 128-column row would not fit one lane's registers.  Lanes 2p and 2p + 1 of wave v
// of workgroup w hold panel row w * 32 NW + 32 v + p, the even and the odd columns of it
// (lane parity h: columns 16 B + 2 e + h, 64 values).  Column i lives in the lanes of parity
// i & 1; the multiplier formed there reaches the partner lane by a DPP swap, and column i + 1,
// the next arg-max's column, is always in the partner.  Keys are 64 bits (the bits of |x| + 2,
// NaN 1, no candidate 0), so equal keys are equal magnitudes and ties go to the smallest label
// exactly as AMax's first index; the waves' and the slots' winners are reduced from small LDS
// tables.  A slot holds the key, the packed label and the row's 128 values as two tagged words
// each; the best row a wave reads goes to LDS split by column parity, so a lane's reads of the
// pivot row are contiguous.  With a single workgroup (m <= 32 NW) nothing is published.
typedef double d8v __attribute__((ext_vector_type(8)));
typedef double d2v __attribute__((ext_vector_type(2)));

// slot: word 0 key (high half), 1 key (low half), 2 packed label, 8 + 2c + {0, 1}: value c
// (high, low); 272 words keep slots on 128-byte lines.  At most 64 workgroups, two parities.
constexpr int RS64_WORDS = 272;
constexpr int W64MAX = 64;
static_assert(2 * (size_t)W64MAX * RS64_WORDS * sizeof(u64) <= PANEL_CAND_BYTES,
              "FP64 row-panel slots exceed the candidate area");

// the value of the partner lane (lane ^ 1): DPP quad_perm [1, 0, 3, 2]
__device__ __forceinline__ double partner(double x) {
  const u64 b = __builtin_bit_cast(u64, x);
  const unsigned lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b), 0xB1, 0xf, 0xf, true);
  const unsigned hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), 0xB1, 0xf, 0xf, true);
  return __builtin_bit_cast(double, (static_cast<u64>(hi) << 32) | lo);
}

__device__ __forceinline__ void put64(u64* p, double v, unsigned tag) {
  const u64 b = __builtin_bit_cast(u64, v);
  put(p, static_cast<unsigned>(b >> 32), tag);
  put(p + 1, static_cast<unsigned>(b), tag);
}
__device__ __forceinline__ double join64(u64 hi, u64 lo) {
  return __builtin_bit_cast(double, ((hi >> 32) << 32) | (lo >> 32));
}

template <int NW, int S>
struct RowsPanel64 {
  // uniform over the workgroup (h: per lane)
  int m, nb, W, w, wave, lane, h, row;
  unsigned base, rounds;
  int qwp;  // wrow slot of the previous step's pivot row
  int32_t* piv;
  int32_t* info;
  int info_base;
  PanelSync sy;
  double (*stage)[NB_MAX + 2];      // per wave: the raw candidate row, its pending multiplier
  double (*wrow)[NW][2][NB_MAX / 2];  // [parity][wave][column parity][64]: best row read
  double (*pval)[NW];               // [parity][wave]: its column-i value
  u64 (*tkey)[NW];                  // [0]: waves' candidates, [1]: waves' best slots
  unsigned (*tlab)[NW];
  int* gaveup;
  // per lane: its half row (blk[B][e] = column 16 B + 2 e + h), label, column-i value (lanes of
  // parity i & 1) and the pending multiplier of the previous step
  d8v blk[NB_MAX / 16];
  int lbl;
  double cur, lprev;

  template <int JB>
  __device__ __forceinline__ void apply_prev(int i) {
    if (i == 0 || !(lbl >= i && lbl != INT_MAX)) return;
    const double nl = -lprev;
    const double* u = wrow[(i - 1) & 1][qwp][h];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const d2v uu = *reinterpret_cast<const d2v*>(u + 8 * JB + 2 * q);
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int k = 2 * q + e;
        const double y = fma(nl, uu[e], blk[JB][k]);
        blk[JB][k] = 16 * JB + 2 * k + h > i ? y : blk[JB][k];
      }
    }
#pragma unroll
    for (int B = JB + 1; B < NB_MAX / 16; ++B) {
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const d2v uu = *reinterpret_cast<const d2v*>(u + 8 * B + 2 * q);
#pragma unroll
        for (int e = 0; e < 2; ++e) blk[B][2 * q + e] = fma(nl, uu[e], blk[B][2 * q + e]);
      }
    }
  }

  // The candidate row (label lab, held by a lane pair of this wave) with its pending update:
  // lane l ends with columns 2l and 2l + 1 (lanes >= 8 JB; zeros when has is false).
  template <int JB>
  __device__ __forceinline__ d2v staged_pair(int i, bool has, int lab) {
    if (has && lbl == lab) {
#pragma unroll
      for (int B = JB; B < NB_MAX / 16; ++B)
#pragma unroll
        for (int e = 0; e < 8; ++e) stage[wave][16 * B + 2 * e + h] = blk[B][e];
      if (h == 0) stage[wave][NB_MAX] = lprev;
    }
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): the staged row
    __builtin_amdgcn_wave_barrier();
    d2v x = {0.0, 0.0};
    if (has && lane >= 8 * JB) {
      const int c0 = 2 * lane;
      x = *reinterpret_cast<const d2v*>(&stage[wave][c0]);
      if (i > 0) {
        const double nl = -stage[wave][NB_MAX];
        const double u0 = wrow[(i - 1) & 1][qwp][0][lane], u1 = wrow[(i - 1) & 1][qwp][1][lane];
        if (c0 > i) x[0] = fma(nl, u0, x[0]);
        if (c0 + 1 > i) x[1] = fma(nl, u1, x[1]);
      }
    }
    return x;
  }

  // the best row of this wave to LDS (columns <= i zeroed) and its column-i value
  __device__ __forceinline__ void keep_row(int i, int par, int slotw, d2v x) {
    const int c0 = 2 * lane;
    wrow[par][slotw][0][lane] = c0 > i ? x[0] : 0.0;
    wrow[par][slotw][1][lane] = c0 + 1 > i ? x[1] : 0.0;
    if (c0 == i) pval[par][slotw] = x[0];
    if (c0 + 1 == i) pval[par][slotw] = x[1];
  }

  template <int JB>
  __device__ __forceinline__ bool step(int ii) {
    const int i = 16 * JB + ii;
    const int par = i & 1;
    const unsigned tag = base + static_cast<unsigned>(i) + 1u;
    const int hi_ = i & 1;  // parity of the lanes holding column i
    // 1. the workgroup's candidate: the wave's winner, then the waves' table
    {
      const bool cand = h == hi_ && lbl >= i && lbl != INT_MAX;
      u64 kw;
      const int wl = lanes_winner64<64>(arg_key64(cur, cand), static_cast<unsigned>(lbl), lane, kw);
      if (lane == 0) {
        tkey[0][wave] = kw;
        tlab[0][wave] = wl >= 0 ? static_cast<unsigned>(__builtin_amdgcn_readlane(lbl, wl))
                                : 0xFFFFFFFFu;
      }
    }
    __syncthreads();
    u64 kwg;
    int labwg;
    {
      const u64 k = lane < NW ? tkey[0][lane] : 0ull;
      const unsigned t = lane < NW ? tlab[0][lane] : 0xFFFFFFFFu;
      const int wq = lanes_winner64<NW>(k, t, lane, kwg);
      labwg = wq >= 0 ? static_cast<int>(__builtin_amdgcn_readlane(t, wq)) : INT_MAX;
    }
    const bool has = kwg != 0ull;
    const bool holder = __ballot(has && lbl == labwg) != 0;
    bool gave = false;
    const int nv = W > 1 && wave < W ? (W - wave + NW - 1) / NW : 0;
    const bool rl = lane >= 8 * JB;
    const u64* area = sy.cand + (size_t)par * W64MAX * RS64_WORDS;
    u64 kw0 = 0, kw1 = 0, kw2 = 0, va[S][4];
    auto issue = [&]() {
      const u64* ks = area + (size_t)(wave + NW * min(lane, nv - 1)) * RS64_WORDS;
      kw0 = get(ks);
      kw1 = get(ks + 1);
      kw2 = get(ks + 2);
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const u64* sl = area + (size_t)(wave + NW * min(j, nv - 1)) * RS64_WORDS;
        const u64* vs = rl ? sl + 8 + 4 * lane : sl;
#pragma unroll
        for (int e = 0; e < 4; ++e) va[j][e] = get(vs + (rl ? e : 0));
      }
    };
    // 2. publish (several workgroups) or stage (one)
    if (W > 1) {
      if (holder || (!has && wave == 0)) {
        const d2v x = staged_pair<JB>(i, has, labwg);
        u64* slot = sy.cand + ((size_t)par * W64MAX + w) * RS64_WORDS;
        if (rl) {
          put64(slot + 8 + 4 * lane, x[0], tag);
          put64(slot + 10 + 4 * lane, x[1], tag);
        }
        if (lane == 0) {
          put(slot, static_cast<unsigned>(kwg >> 32), tag);
          put(slot + 1, static_cast<unsigned>(kwg), tag);
          put(slot + 2, static_cast<unsigned>(has ? labwg * 256 + w : INT_MAX), tag);
        }
      }
      // 3. the first reads of this wave's slots go out before the pending update
      if (nv > 0) issue();
    } else {
      if (holder) keep_row(i, par, 0, staged_pair<JB>(i, true, labwg));
      if (wave == 0 && lane == 0) {
        tkey[1][0] = kwg;
        tlab[1][0] = has ? static_cast<unsigned>(labwg) * 256u : 0xFFFFFFFFu;
        for (int v = 1; v < NW; ++v) tkey[1][v] = 0ull;
      }
    }
    apply_prev<JB>(i);
    if (nv > 0) {
      for (;;) {
        bool ok = lane >= nv || (static_cast<unsigned>(kw0) == tag &&
                                 static_cast<unsigned>(kw1) == tag &&
                                 static_cast<unsigned>(kw2) == tag);
#pragma unroll
        for (int j = 0; j < S; ++j)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            ok = ok && (!rl || j >= nv || static_cast<unsigned>(va[j][e]) == tag);
        if (__all(ok)) break;  // wave-uniform loop: the round counter stays scalar
        if (!keep_waiting(sy.keys, base + 1u, rounds)) {
          gave = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        issue();
      }
      u64 kb = 0ull;
      unsigned pk = 0xFFFFFFFFu;
      if (!gave) {
        const u64 kj = lane < nv ? (((kw0 >> 32) << 32) | (kw1 >> 32)) : 0ull;
        const unsigned pj = lane < nv ? static_cast<unsigned>(kw2 >> 32) : 0xFFFFFFFFu;
        const int jb = lanes_winner64<S>(kj, pj, lane, kb);
        if (jb >= 0) {
          pk = __builtin_amdgcn_readlane(pj, jb);
          if (rl) {
            u64 a[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              a[e] = va[0][e];
#pragma unroll
              for (int j = 1; j < S; ++j) a[e] = j == jb ? va[j][e] : a[e];
            }
            keep_row(i, par, wave, d2v{join64(a[0], a[1]), join64(a[2], a[3])});
          }
        }
      } else if (lane == 0) {
        *gaveup = 1;
      }
      if (lane == 0) {
        tkey[1][wave] = kb;
        tlab[1][wave] = pk;
      }
    } else if (W > 1 && lane == 0) {
      tkey[1][wave] = 0ull;
      tlab[1][wave] = 0xFFFFFFFFu;
    }
    __syncthreads();
    // 4. the winner, the same in every workgroup
    if (*gaveup) return true;
    u64 gk;
    const int qw = lanes_winner64<NW>(lane < NW ? tkey[1][lane] : 0ull,
                                      lane < NW ? tlab[1][lane] : 0xFFFFFFFFu, lane, gk);
    if (qw < 0) return true;  // no candidate (m >= nb: never)
    const unsigned gp = tlab[1][qw];
    const int p = static_cast<int>(gp >> 8);  // label of the pivot row
    const double pv = pval[par][qw];
    if (w == 0 && wave == 0 && lane == 0) piv[i] = p - i;
    const bool stop = pv == 0.0;
    if (stop && w == 0 && wave == 0 && lane == 0) atomicCAS(info, 0, info_base + i + 1);
    lbl = lbl == i ? p : (lbl == p ? i : lbl);
    qwp = qw;
    if (stop) return true;
    // 5. multiplier (lanes of parity i & 1, shared with the partner), and column i + 1
    constexpr int JN = JB + 1 < NB_MAX / 16 ? JB + 1 : JB;
    const bool act = lbl > i && lbl != INT_MAX;
    const double lown = act && h == hi_ ? cur * (1.0 / pv) : 0.0;
    // the swap runs with every lane active (its source is the other lane of the pair; a swap
    // moved under the branch of the receiving lanes would read inactive lanes as 0)
    const double lpart = opaque(partner(lown));
    const double l = h == hi_ ? lown : lpart;
    if (act) {
      const int ei = ii >> 1;  // column i: element ei of block JB in the lanes of parity hi_
      if (h == hi_) blk[JB][ei] = l;
      lprev = l;
      if (i + 1 < nb && h != hi_) {
        const double un = wrow[par][qw][h][(i + 1) >> 1];
        if (ii < 15) {
          const int en = (ii + 1) >> 1;
          const double y = fma(-l, un, blk[JB][en]);
          blk[JB][en] = y;
          cur = y;
        } else {
          blk[JN][0] = fma(-l, un, blk[JN][0]);
          cur = blk[JN][0];
        }
      }
    } else if (h != hi_) {
      cur = ii < 15 ? blk[JB][(ii + 1) >> 1] : blk[JN][0];
    }
    return false;
  }

  template <int JB>
  __device__ __forceinline__ bool block() {
    const int ie = min(16, nb - 16 * JB);
    for (int ii = 0; ii < ie; ++ii)
      if (step<JB>(ii)) return true;
    return false;
  }
};

template <int NW, int S>
__global__ __launch_bounds__(NW * 64, 1) void getf2_rows64_kernel(int m, int nb,
                                                                  double* __restrict__ P,
                                                                  int64_t ld,
                                                                  int32_t* __restrict__ piv,
                                                                  int32_t* __restrict__ info,
                                                                  PanelSync sy, unsigned base,
                                                                  int info_base) {
  __shared__ __attribute__((aligned(16))) double stage[NW][NB_MAX + 2];
  __shared__ __attribute__((aligned(16))) double wrow[2][NW][2][NB_MAX / 2];
  __shared__ double pval[2][NW];
  __shared__ u64 tkey[2][NW];
  __shared__ unsigned tlab[2][NW];
  __shared__ int gaveup;

  RowsPanel64<NW, S> st;
  st.m = m;
  st.nb = nb;
  st.W = gridDim.x;
  st.w = blockIdx.x;
  const int tid = threadIdx.x;
  st.lane = tid & 63;
  st.h = st.lane & 1;
  st.wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  st.row = st.w * 32 * NW + 32 * st.wave + (st.lane >> 1);
  st.base = base;
  st.rounds = 0;
  st.qwp = 0;
  st.piv = piv;
  st.info = info;
  st.info_base = info_base;
  st.sy = sy;
  st.stage = stage;
  st.wrow = wrow;
  st.pval = pval;
  st.tkey = tkey;
  st.tlab = tlab;
  st.gaveup = &gaveup;

  if (st.w == 0 && tid == 0) {
    // ordered before step 0's publication, which every workgroup waits for before its
    // final atomicAdd on the counter
    __hip_atomic_store(sy.list_count, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __atomic_thread_fence(__ATOMIC_RELEASE);
  }
  // an earlier panel already hit a zero pivot: leave everything as it is (empty list)
  if (__builtin_amdgcn_readfirstlane(
          __hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0)
    return;
  if (tid == 0) gaveup = 0;

  const int row = st.row, h = st.h;
  const bool in = row < m;
  st.lbl = in ? row : INT_MAX;
  st.lprev = 0.0;
  const double* prow = P + (int64_t)(in ? row : 0) * ld;
#pragma unroll
  for (int B = 0; B < NB_MAX / 16; ++B)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int c = 16 * B + 2 * e + h;
      st.blk[B][e] = (in && c < nb) ? prow[c] : 0.0;
    }
  st.cur = st.blk[0][0];  // column 0 (lanes of parity 0)
  __syncthreads();        // gaveup is cleared before any wave uses it

  bool stop = st.template block<0>();
  if (!stop && nb > 16) stop = st.template block<1>();
  if (!stop && nb > 32) stop = st.template block<2>();
  if (!stop && nb > 48) stop = st.template block<3>();
  if (!stop && nb > 64) stop = st.template block<4>();
  if (!stop && nb > 80) stop = st.template block<5>();
  if (!stop && nb > 96) stop = st.template block<6>();
  if (!stop && nb > 112) stop = st.template block<7>();

  // every half row to the position of its label, and the row movement list (dst <- src)
  const int lb = st.lbl;
  if (lb != INT_MAX) {
    double* dst = P + (int64_t)lb * ld;
#pragma unroll
    for (int B = 0; B < NB_MAX / 16; ++B)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int c = 16 * B + 2 * e + h;
        if (c < nb) dst[c] = st.blk[B][e];
      }
    if (h == 0 && lb != row) {
      const int slot = atomicAdd(sy.list_count, 1);
      sy.list[2 * slot] = lb;
      sy.list[2 * slot + 1] = row;
    }
  }
}

// Applies the (dst <- src) row list of a panel to columns [0, ncols) of A (rows relative to
// row0), skipping columns [skip_lo, skip_hi).  The moved rows form a permutation among
// themselves, so each workgroup owns LW columns outright: it reads every source row of its
// columns into LDS, then writes them to their destinations, and no other workgroup touches
// those columns.  The list is staged in LDS first; the reads of a thread (one per 64 rows
// moved) are all in flight together.  One launch, no global staging.
template <typename T>
__global__ __launch_bounds__(1024) void laswp_list_kernel(T* __restrict__ A, int64_t ld,
                                                          int64_t row0, int64_t ncols,
                                                          int64_t skip_lo, int64_t skip_hi,
                                                          const int* __restrict__ list,
                                                          const int* __restrict__ count) {
  constexpr int LW = 256 / sizeof(T);  // columns per workgroup: 256-byte row segments
  constexpr int RP = 1024 / LW;        // rows per pass
  constexpr int MAXE = 2 * NB_MAX;     // moved rows of one panel
  __shared__ T buf[MAXE][LW];
  __shared__ int lst[2 * MAXE];
  const int tid = threadIdx.x;
  const int cnt = min(*count, MAXE);
  if (tid < 2 * cnt) lst[tid] = list[tid];
  __syncthreads();
  const int64_t width = ncols - (skip_hi - skip_lo);
  const int lc = tid % LW, rs = tid / LW;
  const int64_t t = (int64_t)blockIdx.x * LW + lc;
  const bool ok = t < width;
  const int64_t c = t < skip_lo ? t : t + (skip_hi - skip_lo);
  T v[MAXE / RP];
#pragma unroll
  for (int q = 0; q < MAXE / RP; ++q) {  // branch-free: all reads in flight together
    const int e = rs + q * RP;
    const bool in = ok && e < cnt;
    v[q] = A[in ? (row0 + lst[2 * e + 1]) * ld + c : row0 * ld];
  }
#pragma unroll
  for (int q = 0; q < MAXE / RP; ++q) {
    const int e = rs + q * RP;
    if (ok && e < cnt) buf[e][lc] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXE / RP; ++q) {
    const int e = rs + q * RP;
    if (ok && e < cnt) A[(row0 + lst[2 * e]) * ld + c] = buf[e][lc];
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
  // The lists are complete before the launch: panel k + 1's count and entries are requested
  // while panel k is applied (entry e < 2 nb lies inside the panel's 4 nb-word list area).
  int ncnt = 0, nd[2] = {0, 0}, ns[2] = {0, 0};
  auto fetch = [&](int k) {
    ncnt = counts[k];
    const int* L = lists + (size_t)k * 4 * nb;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = threadIdx.x + 256 * q;
      if (e < 2 * nb) {
        nd[q] = L[2 * e];
        ns[q] = L[2 * e + 1];
      }
    }
  };
  if (npanels > 0) fetch(0);
  __syncthreads();
  for (int k = 0; k < npanels; ++k) {
    const int row0 = k * nb;
    const int cnt = ncnt;
    const int d[2] = {nd[0], nd[1]}, sr[2] = {ns[0], ns[1]};
    if (k + 1 < npanels) fetch(k + 1);
    int v[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = threadIdx.x + 256 * q;
      v[q] = (e < cnt) ? perm[row0 + sr[q]] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = threadIdx.x + 256 * q;
      if (e < cnt) perm[row0 + d[q]] = v[q];
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += 256) perm_out[i] = perm[i];
}

// The row moves of one block of PB consecutive relative pivots (rows i and i + piv[i]
// exchanged in order of i; piv[i] <= 0, or a partner past n, exchanges nothing) in the list form
// perm_compose_kernel applies: one wave replays the block's exchanges on a table of the
// positions they touch -- the block's own rows directly, later rows in an extra list searched
// by a ballot -- and writes (dst, src) pairs relative to the block's first row: after the
// block, position dst holds what position src held before it.
constexpr int PB = 128;
__global__ __launch_bounds__(64) void piv_block_lists_kernel(int n, const int32_t* __restrict__ piv,
                                                             int* __restrict__ counts,
                                                             int* __restrict__ lists) {
  __shared__ int md[PB];             // content of the block's rows
  __shared__ int ep[PB], ev[PB];     // later rows touched: position, content
  const int k = blockIdx.x, b0 = k * PB, rows = min(PB, n - b0), lane = threadIdx.x;
  for (int t = lane; t < PB; t += 64) md[t] = b0 + t;
  const int p0 = lane < rows ? piv[b0 + lane] : 0;
  const int p1 = lane + 64 < rows ? piv[b0 + lane + 64] : 0;
  __syncthreads();
  int m = 0;  // extra rows so far (the same in every lane)
  for (int r = 0; r < rows; ++r) {
    const int p = __shfl(r < 64 ? p0 : p1, r & 63);
    const int t = b0 + r + p;
    if (p <= 0 || t >= n) continue;  // uniform
    if (t < b0 + rows) {
      if (lane == 0) {
        const int a = md[r];
        md[r] = md[t - b0];
        md[t - b0] = a;
      }
    } else {
      const unsigned long long h0 = __ballot(lane < m && ep[lane] == t);
      const unsigned long long h1 = __ballot(lane + 64 < m && ep[lane + 64] == t);
      const int j = h0 ? __ffsll((long long)h0) - 1 : h1 ? 64 + __ffsll((long long)h1) - 1 : m;
      if (lane == 0) {
        if (j == m) ev[j] = t;
        ep[j] = t;
        const int a = md[r];
        md[r] = ev[j];
        ev[j] = a;
      }
      m += j == m;
    }
    __syncthreads();
  }
  int* L = lists + (size_t)k * 4 * PB;
  for (int e = lane; e < rows + m; e += 64) {
    const bool own = e < rows;
    L[2 * e] = own ? e : ep[e - rows] - b0;
    L[2 * e + 1] = (own ? md[e] : ev[e - rows]) - b0;
  }
  if (lane == 0) counts[k] = rows + m;
}

// *bad = 1 + the first i with piv[i] > 0 and i + piv[i] >= n (a swap outside the matrix, where
// the reference's SwapRow throws), or 0.  One workgroup; every entry is checked.
__global__ void piv_check_kernel(int n, const int32_t* __restrict__ piv, int* __restrict__ bad) {
  __shared__ int first[16];
  int mine = INT_MAX;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int p = piv[i];
    if (p > 0 && (int64_t)i + p >= n) mine = min(mine, i);
  }
  for (int off = 32; off > 0; off >>= 1) mine = min(mine, __shfl_xor(mine, off));
  if ((threadIdx.x & 63) == 0) first[threadIdx.x >> 6] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    int f = INT_MAX;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) f = min(f, first[w]);
    *bad = f == INT_MAX ? 0 : f + 1;
  }
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

// Waves per panel workgroup for R rows of T.  The per-step update is issue-bound on each SIMD,
// so two waves per SIMD (8 per workgroup) overlap one wave's latency chains with the other's
// work; more waves only lengthen the barriers (measured at m = 2048..16384, R = 64..256).
template <typename T, int R>
constexpr int panel_waves() {
  return R >= 64 ? 8 : 4;
}

template <typename T, int R>
const void* panel_kernel() {
  return reinterpret_cast<const void*>(getf2_panel_kernel<T, R, panel_waves<T, R>()>);
}

template <typename T, int R>
int launch_panel_r(int m, int nb, T* P, int64_t ld, int32_t* piv, int32_t* info,
                   const PanelSync& sy, unsigned base, int info_base, hipStream_t s) {
  // The W workgroups wait on each other, so all of them must become resident: W is at most
  // the number that fit on the device at once (fits()), and every other kernel that may run
  // beside the panel finishes without waiting on it, so the workgroups not yet placed get a
  // CU as soon as such a kernel's workgroups retire.  A plain launch suffices; the
  // cooperative launch costs about 30 us of dispatch overhead per panel on top.
  const int W = ceil_div(m, R);
  hipLaunchKernelGGL((getf2_panel_kernel<T, R, panel_waves<T, R>()>), dim3(W),
                     dim3(64 * panel_waves<T, R>()), 0, s, m, nb, P, ld, piv, info, sy, base,
                     info_base);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

// The FP32 row-per-lane panel: W = ceil(m / 512) workgroups of 8 waves, one per CU; wave v
// reads slots v, v + 8, ..., so S (a power of two >= W / 8) slots per wave and W <= 64.
constexpr int ROWS_NW = 8;
constexpr int ROWS_SMAX = 8;
// from this many workgroups on, the two-phase exchange (keys, then the winner's row) beats
// reading every slot's row with its key
constexpr int ROWS_TWO_W = 17;

template <int S, bool TWO>
int launch_rows_s(int W, int m, int nb, float* P, int64_t ld, int32_t* piv, int32_t* info,
                  const PanelSync& sy, unsigned base, int info_base, hipStream_t s) {
  hipLaunchKernelGGL((getf2_rows_kernel<ROWS_NW, S, TWO>), dim3(W), dim3(64 * ROWS_NW), 0, s, m,
                     nb, P, ld, piv, info, sy, base, info_base);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

// DDLA_FP32_PANEL: "labels" the column-per-lane kernel; "rows1" / "rows2" the row-per-lane
// kernel with the one-phase / two-phase exchange at every width; unset: the row-per-lane
// kernel, two-phase exchange from ROWS_TWO_W workgroups on.
int fp32_panel_mode() {
  static int mode = -1;
  if (mode < 0) {
    const char* e = getenv("DDLA_FP32_PANEL");
    mode = !e ? 1 : strcmp(e, "labels") == 0 ? 0 : strcmp(e, "rows1") == 0 ? 2
                  : strcmp(e, "rows2") == 0 ? 3 : 1;
  }
  return mode;
}

int launch_rows_panel(int m, int nb, float* P, int64_t ld, int32_t* piv, int32_t* info,
                      const PanelSync& sy, unsigned base, int info_base, hipStream_t s) {
  const int W = ceil_div(m, 64 * ROWS_NW);
  const int mode = fp32_panel_mode();
  const bool two = W > 1 && (mode == 3 || (mode == 1 && W >= ROWS_TWO_W));
  if (two) return launch_rows_s<1, true>(W, m, nb, P, ld, piv, info, sy, base, info_base, s);
  const int S = ceil_div(W, ROWS_NW);
  if (S <= 1) return launch_rows_s<1, false>(W, m, nb, P, ld, piv, info, sy, base, info_base, s);
  if (S <= 2) return launch_rows_s<2, false>(W, m, nb, P, ld, piv, info, sy, base, info_base, s);
  if (S <= 4) return launch_rows_s<4, false>(W, m, nb, P, ld, piv, info, sy, base, info_base, s);
  return launch_rows_s<8, false>(W, m, nb, P, ld, piv, info, sy, base, info_base, s);
}

// The row-per-lane panel takes FP32 panels of up to 64 workgroups (m <= 32768) that fit on
// the device at once.
bool rows_panel_fits(int64_t m) {
  const int64_t W = (m + 64 * ROWS_NW - 1) / (64 * ROWS_NW);
  return fp32_panel_mode() != 0 && W <= ROWS_NW * ROWS_SMAX && W <= ctx().num_cus;
}

// The FP64 panel with two lanes per row: W = ceil(m / 256) workgroups of 8 waves.  It takes
// only panels of one workgroup (m <= 256), where it is faster than the column-per-lane kernel
// (312 vs 404 us at m = 256); with more workgroups its 2 KB slots double the exchange traffic
// of the FP32 form and it is slower (m = 1024: 439 vs 413 us; 8192: 1033 vs 515 us).  The
// launch handles up to 32 workgroups (4 slots per wave) for measurements.
// DDLA_FP64_PANEL=labels selects the column-per-lane kernel everywhere.
template <int S>
int launch_rows64_s(int W, int m, int nb, double* P, int64_t ld, int32_t* piv, int32_t* info,
                    const PanelSync& sy, unsigned base, int info_base, hipStream_t s) {
  hipLaunchKernelGGL((getf2_rows64_kernel<ROWS_NW, S>), dim3(W), dim3(64 * ROWS_NW), 0, s, m, nb,
                     P, ld, piv, info, sy, base, info_base);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_rows64_panel(int m, int nb, double* P, int64_t ld, int32_t* piv, int32_t* info,
                        const PanelSync& sy, unsigned base, int info_base, hipStream_t s) {
  const int W = ceil_div(m, 32 * ROWS_NW);
  const int S = ceil_div(W, ROWS_NW);
  if (S <= 1) return launch_rows64_s<1>(W, m, nb, P, ld, piv, info, sy, base, info_base, s);
  if (S <= 2) return launch_rows64_s<2>(W, m, nb, P, ld, piv, info, sy, base, info_base, s);
  return launch_rows64_s<4>(W, m, nb, P, ld, piv, info, sy, base, info_base, s);
}

bool rows64_panel_fits(int64_t m) {
  static int mode = -1;
  if (mode < 0) {
    const char* e = getenv("DDLA_FP64_PANEL");
    mode = (e && strcmp(e, "labels") == 0) ? 0 : 1;
  }
  const int64_t W = (m + 32 * ROWS_NW - 1) / (32 * ROWS_NW);
  return mode == 1 && W <= 1;
}

template <typename T, int R>
int max_coresident() {
  static int cached = -1;
  if (cached < 0) {
    int per_cu = 0;
    DDLA_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, panel_kernel<T, R>(), 64 * panel_waves<T, R>(), 0));
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

size_t panel_area_bytes() {
  return PANEL_KEY_BYTES + PANEL_CAND_BYTES + PANEL_LIST_BYTES + PANEL_ABORT_BYTES;
}

unsigned* panel_abort_word(const PanelScratch* ext) {
  void* area = ext ? ext->area : ctx().ws[WS_PANEL];
  if (!area) return nullptr;
  return reinterpret_cast<unsigned*>(static_cast<char*>(area) + PANEL_ABORT_OFF);
}

// exchange area for panel kernels; *base: the launch's tag range is base + 1 .. base + nb + 2
// (a fresh range per launch, so words of earlier launches never match).  ext: an area the
// caller provides (cleared once, its tag range chosen by the caller) instead of the workspace.
int panel_sync(PanelSync* sy, int nb, unsigned* base, hipStream_t s, const PanelScratch* ext) {
  const size_t kb = PANEL_KEY_BYTES, cb = PANEL_CAND_BYTES;
  const size_t bytes = panel_area_bytes();
  void* ws = nullptr;
  if (ext) {
    ws = ext->area;
    *base = ext->first_tag - 1u;
  } else {
    unsigned first = 0;
    DDLA_RETURN_IF(next_tag(WS_PANEL, bytes, static_cast<unsigned>(nb) + 2u, s, &first));
    DDLA_RETURN_IF(workspace(WS_PANEL, bytes, &ws));
    *base = first - 1u;
  }
  char* p = static_cast<char*>(ws);
  sy->keys = reinterpret_cast<u64*>(p);
  sy->cand = reinterpret_cast<u64*>(p + kb);
  sy->list = reinterpret_cast<int*>(p + kb + cb);
  sy->list_count = sy->list + 4 * NB_MAX;
  // a limit other than the default travels in the area (once changed, it is written before
  // every launch: the area may have been cleared since)
  if (ctx().wait_limit_set)
    DDLA_RETURN_IF(launch_fill_u32(p + PANEL_ABORT_OFF + sizeof(unsigned), 1,
                                   ctx().wait_limit, s));
  return 0;
}

template <typename T>
int launch_getf2_panel_ex(int64_t m, int nb, T* P, int64_t ld, int32_t* piv, int32_t* info,
                          int info_base, int* list_out, int* count_out, hipStream_t s,
                          const PanelScratch* ext) {
  if (nb > NB_MAX || nb <= 0 || m < nb) {
    set_error("panel %lld x %d outside 1 <= nb <= min(m, %d)", (long long)m, nb, NB_MAX);
    return DDLA_ERR_INVALID_ARG;
  }
  PanelSync sy;
  unsigned base = 0;
  DDLA_RETURN_IF(panel_sync(&sy, nb, &base, s, ext));
  if (list_out) {
    sy.list = list_out;
    sy.list_count = count_out;
  }
  ProfScope prof(sizeof(T) == 8 ? "getf2_panel_f64" : "getf2_panel_f32",
                 (double)m * nb * nb - (double)nb * nb * nb / 3.0, 2.0 * sizeof(T) * m * nb, s);
  if constexpr (sizeof(T) == 4) {
    if (rows_panel_fits(m))
      return launch_rows_panel((int)m, nb, P, ld, piv, info, sy, base, info_base, s);
  } else {
    if (rows64_panel_fits(m))
      return launch_rows64_panel((int)m, nb, P, ld, piv, info, sy, base, info_base, s);
  }
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
  if (max_count > 2 * NB_MAX) {
    set_error("row list of %d entries exceeds %d", max_count, 2 * NB_MAX);
    return DDLA_ERR_INVALID_ARG;
  }
  // bytes: every moved row (at most max_count) is read and written once
  ProfScope prof("laswp_list", 0.0, 2.0 * sizeof(T) * max_count * width, s);
  constexpr int LW = 256 / sizeof(T);
  hipLaunchKernelGGL((laswp_list_kernel<T>), dim3((unsigned)ceil_div(width, LW)), dim3(1024), 0,
                     s, A, ld, row0, ncols, skip_lo, skip_hi, list, count);
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

// perm[i] = the row that relative pivots piv[0..n) bring to position i: the moves of each block
// of PB pivots in parallel (piv_block_lists_kernel), then the blocks in order
// (perm_compose_kernel), instead of n exchanges in one lane.
int launch_piv_to_perm(int64_t n, const int32_t* piv, int32_t* perm, hipStream_t s) {
  if (n <= 0) return 0;
  if ((size_t)n * sizeof(int) > 160 * 1024) {
    set_error("permutation of %lld rows does not fit in LDS", (long long)n);
    return DDLA_ERR_UNSUPPORTED;
  }
  const int nblk = ceil_div(n, PB);
  void* ws = nullptr;
  DDLA_RETURN_IF(workspace(WS_PIVLIST, (size_t)nblk * (1 + 4 * PB) * sizeof(int), &ws));
  int* counts = static_cast<int*>(ws);
  int* lists = counts + nblk;
  hipLaunchKernelGGL(piv_block_lists_kernel, dim3(nblk), dim3(64), 0, s, (int)n, piv, counts,
                     lists);
  DDLA_HIP_CHECK(hipGetLastError());
  return launch_perm_compose((int)n, nblk, PB, lists, counts, perm, s);
}

__global__ void fill_u32_kernel(unsigned* __restrict__ p, int64_t count, unsigned v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

int launch_fill_u32(void* p, int64_t count, unsigned v, hipStream_t s) {
  if (count <= 0) return 0;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(count)), dim3(256), 0, s,
                     static_cast<unsigned*>(p), count, v);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

int check_pivots(int64_t n, const int32_t* piv, hipStream_t s) {
  if (n <= 0) return 0;
  void* ws = nullptr;
  DDLA_RETURN_IF(workspace(WS_SMALL, 64, &ws));
  int* bad = static_cast<int*>(ws);
  hipLaunchKernelGGL(piv_check_kernel, dim3(1), dim3(1024), 0, s, (int)n, piv, bad);
  DDLA_HIP_CHECK(hipGetLastError());
  int h = 0;
  DDLA_HIP_CHECK(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  DDLA_HIP_CHECK(hipStreamSynchronize(s));
  if (h != 0) {
    set_error("the pivot at index %d swaps outside the %lld rows (SwapRow throws there)", h - 1,
              (long long)n);
    return DDLA_ERR_INVALID_ARG;
  }
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
                                           int, int*, int*, hipStream_t, const PanelScratch*);
template int launch_getf2_panel_ex<float>(int64_t, int, float*, int64_t, int32_t*, int32_t*, int,
                                          int*, int*, hipStream_t, const PanelScratch*);
template int launch_laswp_list<double>(double*, int64_t, int64_t, int64_t, int64_t, int64_t,
                                       const int*, const int*, int, hipStream_t);
template int launch_laswp_list<float>(float*, int64_t, int64_t, int64_t, int64_t, int64_t,
                                      const int*, const int*, int, hipStream_t);
template int launch_gather_rows<double>(int64_t, int64_t, const double*, int64_t, const int32_t*,
                                        double*, int64_t, bool, hipStream_t);
template int launch_gather_rows<float>(int64_t, int64_t, const float*, int64_t, const int32_t*,
                                       float*, int64_t, bool, hipStream_t);

}  // namespace ddla
