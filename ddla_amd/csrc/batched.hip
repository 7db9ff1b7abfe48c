// Batched factorization of many small problems of uneven order (n <= 256) in one persistent
// launch: LU with partial pivoting and lower Cholesky, FP64 or FP32.
//
// No counterpart in the reference; each problem is the factorization DDLA runs one matrix at
// a time (PLUDecUnblock / GetAndApplyPivUnblock, DDLA/Factorizations/LU.cs:344-374, and
// CholeskyLowerUnblock, DDLA/Factorizations/Cholesky.cs:127-140), here in a right-looking
// blocked form with panels of NB columns.
//
// One workgroup (256 threads) owns one problem at a time; thread t holds row t of the current
// panel in registers (n <= 256 rows), so the pivot search is a wave DPP reduction plus one
// LDS exchange, and the panel's rank-1 updates need no memory traffic.  After each panel:
//   LU:  the panel's row interchanges are applied to the other columns by a gather on read
//        (each row reads its source row, a barrier, then every row writes), U12 = L11^-1 A12
//        in LDS, and the trailing rows A22 -= L21 U12 with L21 still in registers;
//   Cholesky: A22 -= L21 L21^T on the lower triangle, L21 shared through LDS.
// The matrix stays in global memory between panels (a problem is at most 512 KB, so it is
// served from L2 / the Infinity cache).
//
// Scheduling: the host sorts the problems by estimated cost and assigns them to the G
// workgroups of the launch longest-first (LPT); the persistent kernel walks its list.  The
// launch is captured once in a hipGraph per plan, and the panel width NB and the number of
// workgroups per CU are chosen by timing the candidates on a copy of the batch.
#include "ddla_internal.hpp"
#include "wave_ops.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <queue>
#include <sstream>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

namespace ddla {
namespace {

constexpr int NT = 256;    // threads per workgroup = maximum order
constexpr int CH = 16;     // column chunk of the updates outside the panel
constexpr int NB_MAX = 32;

// M[idx] if ok, else 0, without a branch: the read always happens (M[0] when !ok), so the
// reads of an unrolled loop are all in flight together instead of one round trip each
template <typename T>
__device__ __forceinline__ T load_or_zero(const T* __restrict__ M, int64_t idx, bool ok) {
  const T x = M[ok ? idx : 0];
  return ok ? x : T(0);
}

template <typename T>
struct BatchArgs {
  const int32_t* n;
  const int32_t* kind;
  const int64_t* off;    // element offset of each problem's matrix
  const int64_t* poff;   // offset of each problem's pivots
  T* a;
  int32_t* piv;
  int32_t* info;
  const int32_t* sched;  // problem indices, grouped by workgroup
  const int32_t* start;  // [G + 1] ranges into sched
};

template <typename T, int NB>
struct Smem {
  unsigned long long sk[4];  // each wave's best key
  int sr[4];
  T prow[NB];            // pivot row of the step (the new row i)
  T srow[NB];            // old row i (moves to the pivot row's place)
  int stop;
  int orig[NT];          // panel-local source row of each panel row after the interchanges
  T dsh;                 // Cholesky: diagonal of the next step
  T lcol[NT];            // Cholesky: column i of the panel's rows
  union {
    T Pn[NT][NB + 1];    // Cholesky: the factored panel (row t of the panel in row t)
    struct {
      T L11[NB][NB + 1];  // LU: the panel's unit lower triangle
      T U12[NB][NT];      // LU: U12 = L11^-1 A12 of all the columns right of the panel
    } lu;
  };
};

// Block-wide arg-max of column entries x over the candidate rows (cand): the largest |x|,
// the smallest row among equal |x|, NaN below every number (AMax picks the first maximum).
// Integer keys (0 not a candidate, 1 NaN, bits(|x|) + 2) reduce by one DPP max per step; ties
// are resolved only when several lanes share the maximum.  Returns the key (<= 2: no nonzero
// candidate) and the row r.
template <typename T, int NB>
__device__ __forceinline__ unsigned long long block_argmax(T x, bool cand, int& r,
                                                           Smem<T, NB>& sm) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long kw = 0;
  int wl;
  if constexpr (sizeof(T) == 8) {
    wl = lanes_winner64<64>(arg_key64(x, cand), (unsigned)threadIdx.x, lane, kw);
  } else {
    unsigned k = 0;
    wl = lanes_winner<64>(arg_key(x, cand), (unsigned)threadIdx.x, lane, k);
    kw = k;
  }
  if (lane == 0) {
    sm.sk[wave] = kw;
    sm.sr[wave] = 64 * wave + max(wl, 0);
  }
  __syncthreads();
  unsigned long long k[4];
  int rr[4];
#pragma unroll
  for (int w = 0; w < 4; ++w) {  // the four entries read at once
    k[w] = sm.sk[w];
    rr[w] = sm.sr[w];
  }
  // waves hold ascending rows: the first wave with the largest key has the smallest row
  unsigned long long kb = k[0];
  r = rr[0];
#pragma unroll
  for (int w = 1; w < 4; ++w) {
    const bool better = k[w] > kb;
    kb = better ? k[w] : kb;
    r = better ? rr[w] : r;
  }
  return kb;
}

// LU with partial pivoting of one n x n row-major problem (ld = n) at M; relative pivots at
// P (LU.cs:386-391); *info = first zero pivot column + 1, and the problem stops there (the
// reference throws from InvScal, BLAS/1v.cs:99).
template <typename T, int NB>
__device__ __forceinline__ void lu_problem(int n, T* __restrict__ M, int32_t* __restrict__ P,
                                           int32_t* __restrict__ info, Smem<T, NB>& sm) {
  const int tid = threadIdx.x;
  if (tid == 0) sm.stop = 0;
  for (int j0 = 0; j0 < n; j0 += NB) {
    const int jb = min(NB, n - j0), m = n - j0;
    const bool own = tid < m;
    T v[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
      v[c] = load_or_zero(M, (int64_t)(j0 + tid) * n + j0 + c, own && c < jb);
    if (own) sm.orig[tid] = tid;
    __syncthreads();
    bool stop = false;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      if (i < jb && !stop) {
        int p = tid;
        const unsigned long long kmax = block_argmax(v[i], own && tid >= i, p, sm);
        if (kmax <= 2ull) {  // every candidate is zero (or NaN): zero pivot
          if (tid == 0) {
            *info = j0 + i + 1;
            sm.stop = 1;
          }
          stop = true;
        } else {
          if (tid == p) {
#pragma unroll
            for (int c = 0; c < NB; ++c) sm.prow[c] = v[c];
          }
          if (tid == i && p != i) {
#pragma unroll
            for (int c = 0; c < NB; ++c) sm.srow[c] = v[c];
          }
          if (tid == 0) {
            P[j0 + i] = p - i;
            const int t = sm.orig[i];
            sm.orig[i] = sm.orig[p];
            sm.orig[p] = t;
          }
          __syncthreads();
          if (p != i) {
            if (tid == i) {
#pragma unroll
              for (int c = 0; c < NB; ++c) v[c] = sm.prow[c];
            } else if (tid == p) {
#pragma unroll
              for (int c = 0; c < NB; ++c) v[c] = sm.srow[c];
            }
          }
          const T rinv = T(1) / sm.prow[i];  // InvScal: multiply by the reciprocal
          if (own && tid > i) {
            const T l = v[i] * rinv;
            v[i] = l;
            // no per-column guard: the columns past jb hold zeros in every row (the pivot
            // row's too) and are never stored, and a guard would serialize one LDS read per
            // column behind its own branch
#pragma unroll
            for (int c = i + 1; c < NB; ++c) v[c] = fma(-l, sm.prow[c], v[c]);
          }
        }
      }
    }
    __syncthreads();
    if (sm.stop) return;  // uniform
    if (own) {
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < jb) M[(int64_t)(j0 + tid) * n + j0 + c] = v[c];
    }
    if (tid < jb) {
#pragma unroll
      for (int c = 0; c < NB; ++c) sm.lu.L11[tid][c] = v[c];
    }
    const int src = own ? sm.orig[tid] : tid;
    const bool moved = own && src != tid;
    // interchanges on the columns left of the panel (gather on read, then write)
    for (int c0 = 0; c0 < j0; c0 += CH) {
      T x[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c)
        x[c] = load_or_zero(M, (int64_t)(j0 + src) * n + c0 + c, moved && c0 + c < j0);
      __syncthreads();
      if (moved) {
#pragma unroll
        for (int c = 0; c < CH; ++c)
          if (c0 + c < j0) M[(int64_t)(j0 + tid) * n + c0 + c] = x[c];
      }
      __syncthreads();
    }
    // Columns right of the panel.  The source row of a panel row t >= jb is t itself or one
    // of rows 0..jb-1 (every interchange involves a row above jb), so:
    //   1. U12 = L11^-1 A12 of every column at once, thread c down column c in registers,
    //      from the source rows of rows 0..jb-1, into LDS;
    //   2. rows t >= jb: A22 -= L21 U12, each thread along its own row (read from its source
    //      row, written to row t) with no barrier: nobody writes rows 0..jb-1 meanwhile;
    //   3. U12 into rows 0..jb-1.
    // Per element the arithmetic is the substitution's and the update's fma chain in k
    // order, as in the unblocked elimination.
    const int w = n - j0 - jb;  // columns right of the panel
    if (w > 0) {
      __syncthreads();  // L11 and the row origins in LDS
      if (tid < w) {
        T u[NB];
#pragma unroll
        for (int r = 0; r < NB; ++r)
          u[r] = load_or_zero(M, (int64_t)(j0 + sm.orig[min(r, jb - 1)]) * n + j0 + jb + tid,
                              r < jb);
#pragma unroll
        for (int r = 1; r < NB; ++r) {
          if (r < jb) {
#pragma unroll
            for (int k = 0; k < r; ++k) u[r] = fma(-sm.lu.L11[r][k], u[k], u[r]);
          }
          __builtin_amdgcn_sched_barrier(0);  // one row's LDS reads at a time
        }
#pragma unroll
        for (int r = 0; r < NB; ++r)
          if (r < jb) sm.lu.U12[r][tid] = u[r];
      }
      __syncthreads();
      if (own && tid >= jb) {
        const T* sr = M + (int64_t)(j0 + src) * n + j0 + jb;
        T* dr = M + (int64_t)(j0 + tid) * n + j0 + jb;
        // the row's multipliers, read back (L2-resident) rather than kept in registers
        // through step 1
        T l21[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) l21[k] = load_or_zero(dr - jb, k, k < jb);
        T x[CH], y[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) x[c] = load_or_zero(sr, c, c < w);
        for (int c0 = 0; c0 < w; c0 += CH) {
          // the next chunk's reads go out before this chunk's arithmetic
#pragma unroll
          for (int c = 0; c < CH; ++c) y[c] = load_or_zero(sr, c0 + CH + c, c0 + CH + c < w);
#pragma unroll
          for (int k = 0; k < NB; ++k) {
            if (k < jb) {
              const T l = l21[k];
#pragma unroll
              for (int c = 0; c < CH; ++c) x[c] = fma(-l, sm.lu.U12[k][c0 + c], x[c]);
            }
            __builtin_amdgcn_sched_barrier(0);  // one k's LDS reads at a time
          }
#pragma unroll
          for (int c = 0; c < CH; ++c)
            if (c0 + c < w) dr[c0 + c] = x[c];
#pragma unroll
          for (int c = 0; c < CH; ++c) x[c] = y[c];
        }
      }
      __syncthreads();
      if (tid < w) {
#pragma unroll
        for (int r = 0; r < NB; ++r)
          if (r < jb) M[(int64_t)(j0 + r) * n + j0 + jb + tid] = sm.lu.U12[r][tid];
      }
    }
    __syncthreads();  // L11 / U12 are rewritten by the next panel
  }
}

// Lower Cholesky of one n x n row-major problem (row stride ld): a11 = sqrt(a11), a21 *= 1/a11,
// A22 -= a21 a21^T on the lower triangle (Cholesky.cs:127-140); only the lower triangle is
// written.  *info = first non-positive pivot + 1.
template <typename T, int NB>
__device__ __forceinline__ void chol_problem(int n, T* __restrict__ M, int64_t ld,
                                             int32_t* __restrict__ info, Smem<T, NB>& sm) {
  const int tid = threadIdx.x;
  for (int j0 = 0; j0 < n; j0 += NB) {
    const int jb = min(NB, n - j0), m = n - j0;
    const bool own = tid < m;
    T v[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
      v[c] = load_or_zero(M, (int64_t)(j0 + tid) * ld + j0 + c, own && c < jb);
    if (tid == 0) sm.dsh = v[0];
    __syncthreads();
    bool stop = false;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      if (i < jb && !stop) {
        const T d = sm.dsh;
        if (!(d > T(0))) {
          if (tid == 0) *info = j0 + i + 1;
          stop = true;
        } else {
          const T s = sqrt(d);
          if (tid == i) v[i] = s;
          if (own && tid > i) {
            if (s != T(1)) v[i] *= T(1) / s;
            if (tid < jb) sm.lcol[tid] = v[i];
          }
          __syncthreads();
          if (own && tid > i) {
            const T xr = -v[i];
            // unguarded for the same reason as in lu_problem: the entries past jb and above
            // the diagonal (c > tid) are never stored or read again
#pragma unroll
            for (int c = i + 1; c < NB; ++c) v[c] = fma(sm.lcol[c], xr, v[c]);
          }
          if (i + 1 < NB && tid == i + 1) sm.dsh = v[min(i + 1, NB - 1)];
          __syncthreads();
        }
      }
    }
    if (stop) return;  // uniform: every thread saw the same pivot
    if (own) {
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < jb && c <= tid) M[(int64_t)(j0 + tid) * ld + j0 + c] = v[c];
#pragma unroll
      for (int c = 0; c < NB; ++c) sm.Pn[tid][c] = v[c];
    }
    __syncthreads();
    // A22 -= L21 L21^T, lower triangle; row tid of the panel is row j0 + tid
    if (own && tid >= jb) {
      const int row = j0 + tid;
      T x[CH], y[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c)
        x[c] = load_or_zero(M, (int64_t)row * ld + j0 + jb + c, j0 + jb + c <= row);
      for (int c0 = j0 + jb; c0 <= row; c0 += CH) {
        // the next chunk's reads go out before this chunk's arithmetic
#pragma unroll
        for (int c = 0; c < CH; ++c)
          y[c] = load_or_zero(M, (int64_t)row * ld + c0 + CH + c, c0 + CH + c <= row);
        const int lc0 = c0 - j0;  // panel row of column c0
#pragma unroll 2
        for (int k = 0; k < jb; ++k) {
          const T xr = -sm.Pn[tid][k];
#pragma unroll
          for (int c = 0; c < CH; ++c)  // rows past the panel read row NT-1 (never stored)
            x[c] = fma(sm.Pn[min(lc0 + c, NT - 1)][k], xr, x[c]);
        }
#pragma unroll
        for (int c = 0; c < CH; ++c)
          if (c0 + c <= row) M[(int64_t)row * ld + c0 + c] = x[c];
#pragma unroll
        for (int c = 0; c < CH; ++c) x[c] = y[c];
      }
    }
    __syncthreads();
  }
}

template <typename T, int NB>
__global__ __launch_bounds__(NT) void batch_kernel(BatchArgs<T> args) {
  __shared__ Smem<T, NB> sm;
  const int b = args.start[blockIdx.x], e = args.start[blockIdx.x + 1];
  for (int t = b; t < e; ++t) {
    const int p = args.sched[t];
    const int n = args.n[p];
    T* M = args.a + args.off[p];
    if (threadIdx.x == 0) args.info[p] = 0;
    if (args.kind[p] == DDLA_BATCH_LU)
      lu_problem<T, NB>(n, M, args.piv + args.poff[p], args.info + p, sm);
    else
      chol_problem<T, NB>(n, M, n, args.info + p, sm);
    __syncthreads();  // the next problem reuses the LDS
  }
}

template <typename T, int NB>
const void* batch_fn() {
  return reinterpret_cast<const void*>(batch_kernel<T, NB>);
}

const void* batch_kernel_for(int prec, int nb) {
  if (prec == DDLA_PREC_FP64) return nb == 16 ? batch_fn<double, 16>() : batch_fn<double, 32>();
  return nb == 16 ? batch_fn<float, 16>() : batch_fn<float, 32>();
}

// relative cost of one problem (used only to order and balance the schedule)
double problem_cost(int n) { return (double)n * 8.0 + (double)n * n * n / 1024.0; }

// LPT: problems by decreasing cost, each to the workgroup with the least work so far
void make_schedule(const std::vector<int32_t>& n, int G, std::vector<int32_t>& sched,
                   std::vector<int32_t>& start) {
  const int64_t count = (int64_t)n.size();
  std::vector<int32_t> order(count);
  for (int64_t i = 0; i < count; ++i) order[i] = (int32_t)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return n[a] > n[b]; });
  typedef std::pair<double, int> Load;
  std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
  for (int g = 0; g < G; ++g) heap.push({0.0, g});
  std::vector<std::vector<int32_t>> lists(G);
  for (int32_t p : order) {
    Load l = heap.top();
    heap.pop();
    lists[l.second].push_back(p);
    l.first += problem_cost(n[p]);
    heap.push(l);
  }
  sched.clear();
  start.assign(G + 1, 0);
  for (int g = 0; g < G; ++g) {
    start[g] = (int32_t)sched.size();
    sched.insert(sched.end(), lists[g].begin(), lists[g].end());
  }
  start[G] = (int32_t)sched.size();
}

}  // namespace

struct BatchPlan {
  int prec = 0;
  int64_t count = 0;
  int nb = 32;
  int wg_per_cu = 2;
  int G = 0;
  void* a = nullptr;
  int32_t* piv = nullptr;
  int32_t* info = nullptr;
  std::vector<int32_t> n;
  // device: n, kind, off, poff, sched, start (one allocation)
  void* dmeta = nullptr;
  int32_t *d_n = nullptr, *d_kind = nullptr, *d_sched = nullptr, *d_start = nullptr;
  int64_t *d_off = nullptr, *d_poff = nullptr;
  int64_t total_elems = 0, total_piv = 0;
  double flops = 0.0, bytes = 0.0;
  std::vector<double> tune_ms;  // per candidate, in candidate order (if tuned)
  std::string tune_key;         // the tuning cache's key of this batch's bucket
  bool from_cache = false;      // shape taken from the tuning cache
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

// ---- tuning cache: one JSON file of (bucket key -> chosen shape) entries, one per line:
//   {"key": "<arch>|batch|<f64|f32>|n<max order, next power of 2>|c<log2 count>|<lu|chol|mixed>",
//    "nb": 32, "wg_per_cu": 4, "ms": 16.1}
// at $DDLA_TUNE_CACHE, else $XDG_CACHE_HOME/ddla_amd/tune.json, else
// $HOME/.cache/ddla_amd/tune.json.  A missing or unwritable file only means no cache.
std::string tune_cache_path() {
  if (const char* p = std::getenv("DDLA_TUNE_CACHE")) return p;
  if (const char* x = std::getenv("XDG_CACHE_HOME")) return std::string(x) + "/ddla_amd/tune.json";
  if (const char* h = std::getenv("HOME")) return std::string(h) + "/.cache/ddla_amd/tune.json";
  return "";
}

std::string tune_bucket_key(int prec, const std::vector<int32_t>& n, const int32_t* kind,
                            int64_t count) {
  hipDeviceProp_t prop{};
  std::string arch = "unknown";
  if (hipGetDeviceProperties(&prop, ctx().device) == hipSuccess) {
    arch = prop.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch = arch.substr(0, colon);
  }
  int nmax = 1, lu = 0;
  for (int64_t i = 0; i < count; ++i) {
    nmax = std::max(nmax, (int)n[(size_t)i]);
    lu += kind[i] == DDLA_BATCH_LU;
  }
  int nb2 = 1;
  while (nb2 < nmax) nb2 *= 2;
  int c2 = 0;
  while ((int64_t(1) << (c2 + 1)) <= count) ++c2;
  const char* mix = lu == count ? "lu" : (lu == 0 ? "chol" : "mixed");
  std::ostringstream k;
  k << arch << "|batch|" << (prec == DDLA_PREC_FP64 ? "f64" : "f32") << "|n" << nb2 << "|c" << c2
    << "|" << mix;
  return k.str();
}

struct TuneEntry {
  int nb = 0, wg = 0;
  double ms = 0.0;
};

std::map<std::string, TuneEntry> tune_cache_read(const std::string& path) {
  std::map<std::string, TuneEntry> out;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    const size_t k = line.find("\"key\": \"");
    if (k == std::string::npos) continue;
    const size_t k0 = k + 8, k1 = line.find('"', k0);
    if (k1 == std::string::npos) continue;
    TuneEntry e;
    const size_t nbp = line.find("\"nb\": "), wgp = line.find("\"wg_per_cu\": "),
                 msp = line.find("\"ms\": ");
    if (nbp == std::string::npos || wgp == std::string::npos) continue;
    e.nb = std::atoi(line.c_str() + nbp + 6);
    e.wg = std::atoi(line.c_str() + wgp + 13);
    if (msp != std::string::npos) e.ms = std::atof(line.c_str() + msp + 6);
    if ((e.nb == 16 || e.nb == 32) && e.wg >= 1 && e.wg <= 8) out[line.substr(k0, k1 - k0)] = e;
  }
  return out;
}

void tune_cache_write(const std::string& path, const std::map<std::string, TuneEntry>& m) {
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos) {
    // create the directory chain (best effort)
    std::string dir = path.substr(0, slash), cur;
    std::stringstream ss(dir);
    std::string part;
    if (!dir.empty() && dir[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
      if (part.empty()) continue;
      cur += part + "/";
      (void)mkdir(cur.c_str(), 0755);
    }
  }
  const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
  {
    std::ofstream f(tmp);
    if (!f) return;
    f << "{\"entries\": [\n";
    size_t i = 0;
    for (const auto& [key, e] : m) {
      f << "  {\"key\": \"" << key << "\", \"nb\": " << e.nb << ", \"wg_per_cu\": " << e.wg
        << ", \"ms\": " << e.ms << "}" << (++i < m.size() ? "," : "") << "\n";
    }
    f << "]}\n";
    if (!f) {
      std::remove(tmp.c_str());
      return;
    }
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
}

int max_wg_per_cu(int prec, int nb) {
  static std::map<std::pair<int, int>, int> cache;
  auto key = std::make_pair(prec, nb);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int per_cu = 0;
  DDLA_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel_for(prec, nb),
                                                              NT, 0));
  cache[key] = per_cu;
  return per_cu;
}

// uploads the schedule for G workgroups into the plan's device metadata
int upload_schedule(BatchPlan* p, int G, hipStream_t s) {
  std::vector<int32_t> sched, start;
  make_schedule(p->n, G, sched, start);
  DDLA_HIP_CHECK(hipMemcpyAsync(p->d_sched, sched.data(), sched.size() * sizeof(int32_t),
                                hipMemcpyHostToDevice, s));
  DDLA_HIP_CHECK(hipMemcpyAsync(p->d_start, start.data(), start.size() * sizeof(int32_t),
                                hipMemcpyHostToDevice, s));
  DDLA_HIP_CHECK(hipStreamSynchronize(s));  // the host vectors go out of scope
  return 0;
}

template <typename T>
int launch_batch(const BatchPlan* p, int nb, int G, void* a, int32_t* piv, int32_t* info,
                 hipStream_t s) {
  BatchArgs<T> args{p->d_n, p->d_kind, p->d_off, p->d_poff, static_cast<T*>(a),
                    piv, info, p->d_sched, p->d_start};
  if (nb == 16)
    hipLaunchKernelGGL((batch_kernel<T, 16>), dim3(G), dim3(NT), 0, s, args);
  else
    hipLaunchKernelGGL((batch_kernel<T, 32>), dim3(G), dim3(NT), 0, s, args);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_plan(const BatchPlan* p, int nb, int G, void* a, int32_t* piv, int32_t* info,
                hipStream_t s) {
  if (p->prec == DDLA_PREC_FP64) return launch_batch<double>(p, nb, G, a, piv, info, s);
  return launch_batch<float>(p, nb, G, a, piv, info, s);
}

}  // namespace

int batch_create(BatchPlan** out, int prec, int64_t count, const int32_t* n_host,
                 const int32_t* kind_host, void* a, int32_t* piv, int32_t* info, int tune) {
  DDLA_ARG_CHECK(out != nullptr, "batch: null plan pointer");
  DDLA_ARG_CHECK(prec == DDLA_PREC_FP64 || prec == DDLA_PREC_FP32,
                 "batch: precision must be FP64 or FP32");
  DDLA_ARG_CHECK(count >= 0 && count < INT32_MAX, "batch: count %lld out of range",
                 (long long)count);
  DDLA_ARG_CHECK(count == 0 || (n_host && kind_host && a && piv && info),
                 "batch: null argument");
  hipStream_t s = stream();
  BatchPlan* p = new BatchPlan();
  p->prec = prec;
  p->count = count;
  p->a = a;
  p->piv = piv;
  p->info = info;
  p->n.assign(n_host, n_host + count);
  std::vector<int64_t> off(count), poff(count);
  const size_t esz = prec == DDLA_PREC_FP64 ? 8 : 4;
  for (int64_t i = 0; i < count; ++i) {
    const int ni = n_host[i];
    if (ni < 1 || ni > NT || (kind_host[i] != DDLA_BATCH_LU && kind_host[i] != DDLA_BATCH_CHOL)) {
      delete p;
      set_error("batch: problem %lld has n = %d (must be 1..%d) or kind %d", (long long)i, ni,
                NT, kind_host[i]);
      return DDLA_ERR_INVALID_ARG;
    }
    off[i] = p->total_elems;
    poff[i] = p->total_piv;
    p->total_elems += (int64_t)ni * ni;
    p->total_piv += ni;
    p->flops += (kind_host[i] == DDLA_BATCH_LU ? 2.0 / 3.0 : 1.0 / 3.0) * ni * (double)ni * ni;
    p->bytes += 2.0 * esz * ni * (double)ni;
  }
  const int cus = ctx().num_cus;
  const int Gmax = cus * 8;
  const size_t bytes = count * (4 + 4 + 8 + 8 + 4) + (Gmax + 1) * 4 + 64;
  if (hipMalloc(&p->dmeta, std::max<size_t>(bytes, 64)) != hipSuccess) {
    delete p;
    set_error("batch: device allocation of %zu bytes failed", bytes);
    return DDLA_ERR_ALLOC;
  }
  char* q = static_cast<char*>(p->dmeta);
  p->d_off = reinterpret_cast<int64_t*>(q);
  p->d_poff = p->d_off + count;
  p->d_n = reinterpret_cast<int32_t*>(p->d_poff + count);
  p->d_kind = p->d_n + count;
  p->d_sched = p->d_kind + count;
  p->d_start = p->d_sched + count;
  int st = 0;
  auto fail = [&](int code) {
    if (p->dmeta) (void)hipFree(p->dmeta);
    delete p;
    return code;
  };
  if (count > 0) {
    if (hipMemcpyAsync(p->d_off, off.data(), count * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(p->d_poff, poff.data(), count * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(p->d_n, n_host, count * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(p->d_kind, kind_host, count * 4, hipMemcpyHostToDevice, s) != hipSuccess) {
      set_error("batch: metadata upload failed");
      return fail(DDLA_ERR_HIP);
    }
  }
  // candidates: panel width x workgroups per CU (bounded by occupancy)
  std::vector<std::pair<int, int>> cand;
  for (int nb : {32, 16}) {
    const int occ = max_wg_per_cu(prec, nb);
    if (occ <= 0) continue;
    for (int k : {1, 2, 4, 8})
      if (k <= occ) cand.push_back({nb, k});
  }
  if (cand.empty()) {
    set_error("batch: kernel cannot be resident");
    return fail(DDLA_ERR_UNSUPPORTED);
  }
  p->nb = cand[0].first;
  p->wg_per_cu = std::min(2, cand[0].second);
  for (auto& c : cand)
    if (c.first == 32 && c.second == 2) p->wg_per_cu = 2;
  // tune = 1: the tuning cache's shape for this batch's bucket if it has one, else time the
  // candidates and store the winner; tune = 2: time and store even on a hit
  if (tune && count > 0) {
    p->tune_key = tune_bucket_key(prec, p->n, kind_host, count);
    const std::string path = tune_cache_path();
    if (tune == 1 && !path.empty()) {
      const auto m = tune_cache_read(path);
      auto it = m.find(p->tune_key);
      if (it != m.end()) {
        for (auto& c : cand)
          if (c.first == it->second.nb && c.second == it->second.wg) {
            p->nb = c.first;
            p->wg_per_cu = c.second;
            p->from_cache = true;
          }
      }
    }
  }
  if (tune && count > 0 && !p->from_cache) {
    // time each candidate on a copy of the batch (the input stays untouched)
    void* ta = nullptr;
    int32_t *tp = nullptr, *ti = nullptr;
    if (hipMalloc(&ta, p->total_elems * esz) != hipSuccess ||
        hipMalloc(&tp, p->total_piv * 4) != hipSuccess || hipMalloc(&ti, count * 4) != hipSuccess) {
      if (ta) (void)hipFree(ta);
      if (tp) (void)hipFree(tp);
      set_error("batch: autotune scratch allocation failed");
      return fail(DDLA_ERR_ALLOC);
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    double best = 1e300;
    for (auto& c : cand) {
      const int G = std::min<int64_t>(cus * c.second, count);
      st = upload_schedule(p, G, s);
      double ms_best = 1e300;
      for (int rep = 0; rep < 2 && st == 0; ++rep) {
        if (hipMemcpyAsync(ta, a, p->total_elems * esz, hipMemcpyDeviceToDevice, s) != hipSuccess) {
          st = DDLA_ERR_HIP;
          break;
        }
        (void)hipEventRecord(e0, s);
        st = launch_plan(p, c.first, G, ta, tp, ti, s);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms_best = std::min<double>(ms_best, ms);
      }
      if (st) break;
      p->tune_ms.push_back(ms_best);
      if (ms_best < best) {
        best = ms_best;
        p->nb = c.first;
        p->wg_per_cu = c.second;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(ta);
    (void)hipFree(tp);
    (void)hipFree(ti);
    if (st) return fail(st);
    const std::string path = tune_cache_path();
    if (!path.empty()) {
      auto m = tune_cache_read(path);
      m[p->tune_key] = TuneEntry{p->nb, p->wg_per_cu, best};
      tune_cache_write(path, m);
    }
  }
  p->G = (int)std::max<int64_t>(1, std::min<int64_t>(cus * p->wg_per_cu, count));
  if (count > 0) {
    st = upload_schedule(p, p->G, s);
    if (st) return fail(st);
    // one persistent launch, captured once
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      set_error("batch: stream capture failed to start");
      return fail(DDLA_ERR_HIP);
    }
    st = launch_plan(p, p->nb, p->G, a, piv, info, s);
    hipError_t e = hipStreamEndCapture(s, &p->graph);
    if (st || e != hipSuccess) {
      set_error("batch: stream capture failed");
      return fail(DDLA_ERR_HIP);
    }
    if (hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(p->graph);
      set_error("batch: graph instantiation failed");
      return fail(DDLA_ERR_HIP);
    }
  }
  *out = p;
  return 0;
}

int batch_execute(BatchPlan* p) {
  DDLA_ARG_CHECK(p != nullptr, "batch: null plan");
  if (p->count == 0) return 0;
  ProfScope prof(p->prec == DDLA_PREC_FP64 ? "batch_factor_f64" : "batch_factor_f32", p->flops,
                 p->bytes, stream());
  DDLA_HIP_CHECK(hipGraphLaunch(p->exec, stream()));
  return 0;
}

int batch_destroy(BatchPlan* p) {
  if (!p) return 0;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->dmeta) {
    (void)hipStreamSynchronize(stream());
    (void)hipFree(p->dmeta);
  }
  delete p;
  return 0;
}

int batch_describe(const BatchPlan* p, char* buf, size_t len) {
  DDLA_ARG_CHECK(p != nullptr && buf != nullptr && len > 0, "batch: null argument");
  std::string t;
  for (size_t i = 0; i < p->tune_ms.size(); ++i) {
    char b[32];
    snprintf(b, sizeof b, "%s%.3f", i ? "," : "", p->tune_ms[i]);
    t += b;
  }
  snprintf(buf, len,
           "{\"count\": %lld, \"nb\": %d, \"wg_per_cu\": %d, \"workgroups\": %d, "
           "\"flops\": %.6g, \"bytes\": %.6g, \"tune_ms\": [%s], \"tune_key\": \"%s\", "
           "\"from_cache\": %s}",
           (long long)p->count, p->nb, p->wg_per_cu, p->G, p->flops, p->bytes, t.c_str(),
           p->tune_key.c_str(), p->from_cache ? "true" : "false");
  return 0;
}

}  // namespace ddla
