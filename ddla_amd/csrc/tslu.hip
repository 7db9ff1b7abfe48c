// Panel factorization with tournament pivoting (FP32), the pivot rule of the mixed-precision
// LU when DDLA_PIVOT_TOURNAMENT is chosen.
//
// No counterpart in the reference, which pivots column by column (GetAndApplyPivUnblock,
// DDLA/Factorizations/LU.cs:344-374).  The method is the tournament pivoting of CALU (Grigori,
// Demmel and Xiang, "CALU: a communication optimal LU factorization algorithm", SIAM J. Matrix
// Anal. Appl. 32 (2011)): the nb pivot rows of an m x nb panel are chosen by a reduction tree
// instead of nb global arg-max steps.
//   * leaves: the panel is cut into stacks of 256 rows; each workgroup runs Gaussian
//     elimination with partial pivoting (GEPP) on its stack in LDS and keeps the nb rows it
//     chose as pivots (their panel row numbers);
//   * tree: two sibling candidate sets (2 nb original rows) are stacked and reduced by GEPP to
//     nb candidates again; the second workgroup to arrive at a node merges (an arrival word per
//     node, tagged with the launch), the first one retires;
//   * root: the last GEPP's pivot rows are the panel's pivots and its factored rows are the
//     panel's L11 \ U11.  The root moves the displaced top rows into the places the selected
//     rows leave, writes L11 \ U11 on top and emits the (dst <- src) row list that the rest of
//     the matrix receives (same list format as the partial-pivoting panel);
//   * then L21 = A21 U11^-1 for the rows below (the panel TRSM of the Cholesky kernels).
// Pivot growth stays close to partial pivoting's in practice (the paper's experiments); the
// refinement recovers FP64 accuracy either way, and a zero pivot in U11 reports info.
//
// One column step of the stack GEPP costs one block arg-max (wave DPP reduction, 4 wave
// results through LDS) and one LDS broadcast of the pivot row; columns go in blocks of 16 held
// in registers, and after each block the rows' trailing columns take the block's rank-16
// update from LDS (the loop code stays small, whatever the panel width).
#include "ddla_internal.hpp"
#include "wave_ops.hpp"

#include <algorithm>
#include <climits>

namespace ddla {

namespace {

constexpr int TS_ROWS = 256;        // stack rows = threads per workgroup
constexpr int TS_NB = 128;          // panel width handled by the tournament
constexpr int TS_CB = 16;           // column block of the stack GEPP
constexpr int TS_LD = TS_NB + 1;    // odd LDS row stride: one row per lane, no bank conflicts
constexpr int TS_REST = TS_NB - TS_CB;

struct TsluArgs {
  int m;              // panel rows (>= TS_NB)
  int leaves;         // ceil(m / TS_ROWS)
  float* P;           // m x TS_NB panel, row-major
  int64_t ld;
  int* sets;          // [nodes][TS_NB] candidate panel rows of each tree node
  int* set_count;     // [nodes]
  unsigned* flags;    // [nodes] arrival words (tag of the last launch that arrived)
  unsigned tag;       // unique per launch, never 0
  int32_t* info;
  int info_base;
  int* list;          // (dst, src) pairs, panel-relative
  int* list_count;
};

constexpr int TS_U12LD = TS_CB + 4;  // U12 row stride: 16-byte aligned rows

struct TsSmem {
  float S[TS_ROWS][TS_LD];             // the stack
  float4 U12[TS_REST][TS_U12LD / 4];   // one block's U12, column-major: U12[c][i]
  float4 Lb[TS_CB][TS_CB / 4];         // the block's L11 (rows of the block's pivots)
  float4 cand[2][4][TS_CB / 4];        // per wave: its best row's block columns (by parity)
  int4 ckey[2];                        // per wave: its best key (by parity)
  int4 cidx[2];                        // per wave: its best stack row
  int idx[TS_ROWS];                    // panel row of each stack row
  int order[TS_NB];                    // stack rows in pivot order
  int sel[TS_NB];                      // root: panel row selected for position i
  int top[TS_NB];                      // root: 1 if top position p is selected
  int disp[TS_NB];                     // root: displaced top rows (ascending)
  int vac[TS_NB];                      // root: vacated rows below the top (pivot order)
  int wcount[4];
  int zero;                            // first zero-pivot column + 1 of the last GEPP
  int go;
  int nlist;
};

// Pivot keys as integers: |x| + 1 in its bit pattern for a row still active (the order of
// non-negative floats), 0 for NaN (the smallest candidate value), -1 for a row already
// chosen.  The wave maximum is one DPP max per step; the lowest lane holding it comes from one
// ballot, so equal keys go to the lowest row.
__device__ __forceinline__ int pivot_key(float x, bool active) {
  const int k = x != x ? 0 : __builtin_bit_cast(int, fabsf(x)) + 1;
  return active ? k : -1;
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_imax(int x) {
  return max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, CTRL, ROWMASK, 0xf, false));
}

__device__ __forceinline__ int wave_max_lane(int key, int* lane_out) {
  int m = dpp_imax<0x111, 0xf>(key);  // row_shr:1
  m = dpp_imax<0x112, 0xf>(m);        // row_shr:2
  m = dpp_imax<0x114, 0xf>(m);        // row_shr:4
  m = dpp_imax<0x118, 0xf>(m);        // row_shr:8
  m = dpp_imax<0x142, 0xa>(m);        // row_bcast:15
  m = dpp_imax<0x143, 0xc>(m);        // row_bcast:31
  const int M = __builtin_amdgcn_readlane(m, 63);
  *lane_out = __builtin_ctzll(__ballot(key == M));
  return M;
}

__device__ __forceinline__ float f4(const float4& q, int e) {
  return e == 0 ? q.x : e == 1 ? q.y : e == 2 ? q.z : q.w;
}

// stack rows [0, cnt) <- panel rows sm.idx[0, cnt) (original values); 16-byte loads, 16 in
// flight per thread
__device__ void load_stack(const TsluArgs& a, int cnt, TsSmem& sm, bool vec) {
  const int tid = threadIdx.x;
  if (vec) {
    const int q = tid % (TS_NB / 4), rb = tid / (TS_NB / 4);  // rows rb, rb + 8, ...
    constexpr int RSTEP = TS_ROWS / (TS_NB / 4);
    constexpr int BATCH = 16;
    for (int k0 = 0; k0 < TS_ROWS / RSTEP; k0 += BATCH) {
      float4 x[BATCH];
#pragma unroll
      for (int k = 0; k < BATCH; ++k) {
        const int r = rb + RSTEP * (k0 + k);
        if (r < cnt)
          x[k] = *reinterpret_cast<const float4*>(a.P + (int64_t)sm.idx[r] * a.ld + 4 * q);
      }
#pragma unroll
      for (int k = 0; k < BATCH; ++k) {
        const int r = rb + RSTEP * (k0 + k);
        if (r < cnt) {
          sm.S[r][4 * q + 0] = x[k].x;
          sm.S[r][4 * q + 1] = x[k].y;
          sm.S[r][4 * q + 2] = x[k].z;
          sm.S[r][4 * q + 3] = x[k].w;
        }
      }
    }
  } else {
    for (int e = tid; e < cnt * TS_NB; e += TS_ROWS) {
      const int r = e / TS_NB, c = e % TS_NB;
      sm.S[r][c] = a.P[(int64_t)sm.idx[r] * a.ld + c];
    }
  }
  __syncthreads();
}

// rows of the previous block's deferred update: x[c] -= sum_i l_i U12[c][i] for the U12 columns
// [u0, u1) of the previous block (column c0 + u), fmaf chains in i order
__device__ __forceinline__ void deferred_update(TsSmem& sm, int row, int c0, int u0, int u1,
                                                const float (&l)[TS_CB]) {
  for (int u = u0; u < u1; ++u) {
    float x = sm.S[row][c0 + u];
#pragma unroll
    for (int e = 0; e < TS_CB / 4; ++e) {
      const float4 q = sm.U12[u][e];
      x = fmaf(-l[4 * e], q.x, x);
      x = fmaf(-l[4 * e + 1], q.y, x);
      x = fmaf(-l[4 * e + 2], q.z, x);
      x = fmaf(-l[4 * e + 3], q.w, x);
    }
    sm.S[row][c0 + u] = x;
  }
}

// GEPP on stack rows [0, cnt) in place (rows are not moved: sm.order lists the pivot rows);
// stops after min(cnt, TS_NB) pivots.  A zero (or NaN) pivot column is recorded in sm.zero
// and the selection goes on with the first remaining row and no elimination.
// One barrier per column: every wave posts its best (key, row) and that row's block columns;
// after the barrier every thread picks the same winner among the four and reads its row.
// Columns go in blocks of 16.  At the end of a block, U12 = L11^-1 A12 is formed for all later
// columns and only the next block's 16 columns take the rank-16 update; the update of the
// columns after those is spread over the next block's 16 column steps, between issuing an LDS
// read and using it, so it fills the latency of the pivot search.
__device__ void stack_gepp(int cnt, TsSmem& sm) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  bool active = tid < cnt;
  const int npiv = min(cnt, TS_NB);
  int zero = 0;           // uniform
  float lp[TS_CB];        // multipliers of the previous block (deferred update)
  bool dact = false;      // this row takes the previous block's deferred update
  int dper = 0;           // deferred U12 columns per column step
#pragma unroll
  for (int k = 0; k < TS_CB; ++k) lp[k] = 0.f;
  for (int j0 = 0; j0 < npiv; j0 += TS_CB) {
    float v[TS_CB];
#pragma unroll
    for (int c = 0; c < TS_CB; ++c) v[c] = sm.S[tid][j0 + c];
#pragma unroll
    for (int i = 0; i < TS_CB; ++i) {
      if (j0 + i < npiv) {  // uniform
        const int par = i & 1;
        int bl;
        const int key = wave_max_lane(pivot_key(v[i], active), &bl);
        if (lane == bl) {
#pragma unroll
          for (int e = 0; e < TS_CB / 4; ++e)
            sm.cand[par][wave][e] = make_float4(v[4 * e], v[4 * e + 1], v[4 * e + 2], v[4 * e + 3]);
        }
        if (lane == 0) {
          reinterpret_cast<int*>(&sm.ckey[par])[wave] = key;
          reinterpret_cast<int*>(&sm.cidx[par])[wave] = wave * 64 + bl;
        }
        __syncthreads();
        const int4 kk = sm.ckey[par];
        const int4 ii = sm.cidx[par];
        // the previous block's update of the columns after this block (independent work)
        if (dact) deferred_update(sm, tid, j0, TS_CB + i * dper, TS_CB + (i + 1) * dper, lp);
        // the wave candidates in wave order: a later wave wins only with a larger key
        int best = kk.x, bw = 0;
        if (kk.y > best) { best = kk.y; bw = 1; }
        if (kk.z > best) { best = kk.z; bw = 2; }
        if (kk.w > best) { best = kk.w; bw = 3; }
        const int bi = bw == 0 ? ii.x : bw == 1 ? ii.y : bw == 2 ? ii.z : ii.w;
        if (tid == bi) active = false;
        if (tid == 0) sm.order[j0 + i] = bi;
        if (best <= 1 && zero == 0) zero = j0 + i + 1;  // largest |x| is 0 (or NaN)
        float prow[TS_CB];
#pragma unroll
        for (int e = 0; e < TS_CB / 4; ++e) {
          const float4 q = sm.cand[par][bw][e];
          prow[4 * e] = q.x;
          prow[4 * e + 1] = q.y;
          prow[4 * e + 2] = q.z;
          prow[4 * e + 3] = q.w;
        }
        const float piv = prow[i];
        const float rinv = piv != 0.f ? 1.f / piv : 0.f;  // multiply by the reciprocal
        if (active) {
          const float l = v[i] * rinv;
          v[i] = l;
#pragma unroll
          for (int c = i + 1; c < TS_CB; ++c) v[c] = fmaf(-l, prow[c], v[c]);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < TS_CB; ++c) sm.S[tid][j0 + c] = v[c];
    __syncthreads();
    const int c1 = j0 + TS_CB;
    if (c1 >= npiv) break;  // uniform: no pivots left to choose
    const int rest = TS_NB - c1;
    {  // the block's L11: row i = the block columns of pivot row i
      const int i = tid / TS_CB, k = tid % TS_CB;
      reinterpret_cast<float*>(&sm.Lb[i][0])[k] = sm.S[sm.order[j0 + i]][j0 + k];
    }
    __syncthreads();
    // U12 = L11^-1 A12 over the block's 16 pivot rows (one column per thread), written into
    // the pivot rows and kept column-major for the updates
    if (tid < rest) {
      const int c = c1 + tid;
      int pr[TS_CB];
#pragma unroll
      for (int e = 0; e < TS_CB / 4; ++e) {
        const int4 o = *reinterpret_cast<const int4*>(&sm.order[j0 + 4 * e]);
        pr[4 * e] = o.x;
        pr[4 * e + 1] = o.y;
        pr[4 * e + 2] = o.z;
        pr[4 * e + 3] = o.w;
      }
      float u[TS_CB];
#pragma unroll
      for (int i = 0; i < TS_CB; ++i) {
        float x = sm.S[pr[i]][c];
#pragma unroll
        for (int k = 0; k < i; ++k) x = fmaf(-f4(sm.Lb[i][k / 4], k % 4), u[k], x);
        u[i] = x;
        sm.S[pr[i]][c] = x;
      }
#pragma unroll
      for (int e = 0; e < TS_CB / 4; ++e)
        sm.U12[tid][e] = make_float4(u[4 * e], u[4 * e + 1], u[4 * e + 2], u[4 * e + 3]);
    }
    __syncthreads();
    // the next block's columns now; the columns after them during the next block
    if (active) deferred_update(sm, tid, c1, 0, TS_CB, v);
#pragma unroll
    for (int k = 0; k < TS_CB; ++k) lp[k] = v[k];
    dact = active;
    dper = (rest - TS_CB) / TS_CB;
    __syncthreads();
  }
  if (tid == 0) sm.zero = zero;
  __syncthreads();
}

// rank of this thread's flag among the flags of threads [0, 128) (two waves), and the total
__device__ __forceinline__ int rank128(bool flag, TsSmem& sm, int* total) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const unsigned long long mask = __ballot(flag);
  const int below = __popcll(mask & ((1ull << lane) - 1ull));
  if (lane == 0 && wave < 2) sm.wcount[wave] = __popcll(mask);
  __syncthreads();
  *total = sm.wcount[0] + sm.wcount[1];
  return below + (wave == 1 ? sm.wcount[0] : 0);
}

// Root: interchanges on the panel's own columns, L11 \ U11 on top, the row list and info.
__device__ void finalize(const TsluArgs& a, TsSmem& sm) {
  const int tid = threadIdx.x;
  if (tid < TS_NB) {
    sm.sel[tid] = sm.idx[sm.order[tid]];
    sm.top[tid] = 0;
  }
  if (tid == 0) {
    sm.nlist = 0;
    if (sm.zero) atomicCAS(a.info, 0, a.info_base + sm.zero);
  }
  __syncthreads();
  if (tid < TS_NB && sm.sel[tid] < TS_NB) sm.top[sm.sel[tid]] = 1;
  __syncthreads();
  int nd = 0, nv = 0;
  const bool displaced = tid < TS_NB && !sm.top[tid];
  const int rd = rank128(displaced, sm, &nd);
  if (displaced) sm.disp[rd] = tid;
  __syncthreads();  // wcount is reused
  const bool vacated = tid < TS_NB && sm.sel[tid] >= TS_NB;
  const int rv = rank128(vacated, sm, &nv);
  if (vacated) sm.vac[rv] = sm.sel[tid];
  __syncthreads();
  // displaced top rows (original contents) into the rows the selected ones leave
  for (int e = tid; e < nd * TS_NB; e += TS_ROWS) {
    const int r = e / TS_NB, c = e % TS_NB;
    a.P[(int64_t)sm.vac[r] * a.ld + c] = a.P[(int64_t)sm.disp[r] * a.ld + c];
  }
  __syncthreads();  // every displaced row is read before the top rows are written
  for (int e = tid; e < TS_NB * TS_NB; e += TS_ROWS) {
    const int i = e / TS_NB, c = e % TS_NB;
    a.P[(int64_t)i * a.ld + c] = sm.S[sm.order[i]][c];
  }
  if (tid < TS_NB && sm.sel[tid] != tid) {
    const int slot = atomicAdd(&sm.nlist, 1);
    a.list[2 * slot] = tid;
    a.list[2 * slot + 1] = sm.sel[tid];
  }
  if (tid < nd) {
    const int slot = atomicAdd(&sm.nlist, 1);
    a.list[2 * slot] = sm.vac[tid];
    a.list[2 * slot + 1] = sm.disp[tid];
  }
  __syncthreads();
  if (tid == 0) *a.list_count = sm.nlist;
}

__global__ __launch_bounds__(TS_ROWS) void tslu_tree_kernel(TsluArgs a) {
  __shared__ TsSmem sm;
  const int tid = threadIdx.x;
  const bool vec = a.ld % 4 == 0 && (reinterpret_cast<uintptr_t>(a.P) & 15) == 0;
  // leaf: rows [r0, r0 + cnt) of the panel
  const int r0 = blockIdx.x * TS_ROWS;
  int cnt = min(TS_ROWS, a.m - r0);
  if (tid < cnt) sm.idx[tid] = r0 + tid;
  __syncthreads();
  load_stack(a, cnt, sm, vec);
  int node = blockIdx.x, off = 0, width = a.leaves;  // this workgroup's node, level offset, width
  for (;;) {
    stack_gepp(cnt, sm);
    if (width == 1) {
      finalize(a, sm);
      return;
    }
    const int win = min(cnt, TS_NB);
    int winner_row = tid < win ? sm.idx[sm.order[tid]] : 0;
    // climb: a node without a sibling (the last of an odd level) moves up unchanged; a level
    // of width >= 3 has a parent level of width >= 2, so the root is always a merge
    for (;;) {
      if (width < 2) return;  // unreachable (see above); never loop on a one-node level
      const int noff = off + width, nwidth = (width + 1) / 2;
      if ((node ^ 1) >= width) {
        node /= 2;
        off = noff;
        width = nwidth;
        continue;
      }
      // publish this node's candidates, then arrive at the parent
      int* mine = a.sets + (size_t)(off + node) * TS_NB;
      if (tid < win) mine[tid] = winner_row;
      if (tid == 0) a.set_count[off + node] = win;
      __threadfence();
      __syncthreads();
      if (tid == 0) {
        const unsigned old = atomicExch(&a.flags[noff + node / 2], a.tag);
        sm.go = old == a.tag;  // the sibling arrived first: this workgroup merges
      }
      __syncthreads();
      if (!sm.go) return;
      __threadfence();
      const int left = off + (node & ~1), right = left + 1;
      const int c0 = __hip_atomic_load(&a.set_count[left], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
      const int c1 = __hip_atomic_load(&a.set_count[right], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
      if (tid < c0) sm.idx[tid] = a.sets[(size_t)left * TS_NB + tid];
      else if (tid < c0 + c1) sm.idx[tid] = a.sets[(size_t)right * TS_NB + tid - c0];
      cnt = c0 + c1;
      __syncthreads();
      load_stack(a, cnt, sm, vec);
      node /= 2;
      off = noff;
      width = nwidth;
      break;
    }
  }
}

}  // namespace

bool tslu_supported(int64_t m, int nb) { return nb == TS_NB && m >= TS_NB && m <= (1 << 24); }

int launch_tslu_panel(int64_t m, int nb, float* P, int64_t ld, int32_t* info, int info_base,
                      int* list_out, int* count_out, hipStream_t s) {
  if (!tslu_supported(m, nb)) {
    set_error("tournament panel needs nb = %d and m >= nb (got %lld x %d)", TS_NB, (long long)m,
              nb);
    return DDLA_ERR_INVALID_ARG;
  }
  const int leaves = (int)((m + TS_ROWS - 1) / TS_ROWS);
  const size_t nodes = 2 * (size_t)leaves + 32;
  const size_t bytes = nodes * (TS_NB + 2) * sizeof(int);
  void* ws = nullptr;
  unsigned tag = 0;  // arrival words of earlier launches never hold this launch's tag
  DDLA_RETURN_IF(next_tag(WS_TSLU, bytes, 1, s, &tag));
  DDLA_RETURN_IF(workspace(WS_TSLU, bytes, &ws));
  TsluArgs a;
  a.m = (int)m;
  a.leaves = leaves;
  a.P = P;
  a.ld = ld;
  a.sets = static_cast<int*>(ws);
  a.set_count = a.sets + nodes * TS_NB;
  a.flags = reinterpret_cast<unsigned*>(a.set_count + nodes);
  a.tag = tag;
  a.info = info;
  a.info_base = info_base;
  a.list = list_out;
  a.list_count = count_out;
  {
    ProfScope prof("tslu_tree_f32", 2.0 * (2 * leaves - 1) * TS_ROWS * TS_NB * TS_NB / 2.0,
                   4.0 * (double)m * TS_NB + 8.0 * TS_NB * TS_NB, s);
    hipLaunchKernelGGL(tslu_tree_kernel, dim3(leaves), dim3(TS_ROWS), 0, s, a);
    DDLA_HIP_CHECK(hipGetLastError());
  }
  // L21 = A21 U11^-1 below the top block
  DDLA_RETURN_IF(launch_trsm_rlt_panel<float>(TS_NB, m - TS_NB, P, ld, P + TS_NB * ld, ld, s, true));
  return 0;
}

}  // namespace ddla
