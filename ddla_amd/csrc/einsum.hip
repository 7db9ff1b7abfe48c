// Binary tensor contraction C = beta C + alpha contract(expr, A, B) on the FP64 MFMA GEMM
// (MArray.Contract, DDLA/Einsum/MArray.Contract.cs:35-125, :252-405; index classes from
// DDLA/Utilities/IndiceUtils.cs:18-130).
//
// The index analysis follows the reference: a symbol repeated within one tensor is a diagonal
// (one index, strides added); symbols in A and B only are contracted (K), in A and C only the
// rows (M), in B and C only the columns (N), in all three a batch; symbols of A alone or of B
// alone are summed (batches that add into the same C), symbols of C alone broadcast (batches
// over C).  Where the reference packs block-scatter matrices for its CPU micro-kernel, each
// batch here gathers A and B into contiguous row-major M x K and K x N operands through
// scatter vectors (one offset per row / column), multiplies them on the matrix cores and
// scatters alpha times the product into C.
#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ddla_internal.hpp"

namespace ddla {
namespace {

constexpr int EIN_MAX_DIMS = 16;

struct Idx {
  int64_t len = 1, sa = 0, sb = 0, sc = 0;
};

// symbol -> (length, stride) with repeated symbols merged (IndiceUtils.Diagonal)
int diagonal(const std::string& sym, int nd, const int64_t* len, const int64_t* str,
             std::map<char, std::pair<int64_t, int64_t>>& out, const char* which) {
  if ((int)sym.size() != nd) {
    set_error("The rank of the tensor must match the length of the symbol (%s: rank %d, '%s').",
              which, nd, sym.c_str());
    return DDLA_ERR_INVALID_ARG;
  }
  for (int i = 0; i < nd; ++i) {
    if (len[i] < 0) {
      set_error("negative length in %s", which);
      return DDLA_ERR_INVALID_ARG;
    }
    auto it = out.find(sym[i]);
    if (it != out.end()) {
      if (it->second.first != len[i]) {
        set_error("The character '%c' already exists with different lengths.", sym[i]);
        return DDLA_ERR_INVALID_ARG;
      }
      it->second.second += str[i];
    } else {
      out[sym[i]] = {len[i], str[i]};
    }
  }
  return 0;
}

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && (s[a] == ' ' || s[a] == '\t')) ++a;
  while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t')) --b;
  return s.substr(a, b - a);
}

// offsets of every multi-index of `idx` (last index fastest): off[i] = sum_d i_d stride_d
void offsets(const std::vector<Idx>& idx, int64_t Idx::*stride, std::vector<int64_t>& off) {
  int64_t total = 1;
  for (const Idx& d : idx) total *= d.len;
  off.assign((size_t)total, 0);
  int64_t rep = 1;
  for (int q = (int)idx.size() - 1; q >= 0; --q) {
    const int64_t L = idx[q].len, st = idx[q].*stride;
    for (int64_t i = 0; i < total; ++i) off[(size_t)i] += ((i / rep) % L) * st;
    rep *= L;
  }
}

__global__ __launch_bounds__(256) void gather_kernel(int64_t rows, int64_t cols,
                                                     const double* __restrict__ src, int64_t base,
                                                     const int64_t* __restrict__ roff,
                                                     const int64_t* __restrict__ coff,
                                                     double* __restrict__ dst) {
  const int64_t total = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int64_t r = e / cols, c = e % cols;
    dst[e] = src[base + roff[r] + coff[c]];
  }
}

__global__ __launch_bounds__(256) void scatter_add_kernel(int64_t rows, int64_t cols,
                                                          const double* __restrict__ src,
                                                          double alpha, double* __restrict__ dst,
                                                          int64_t base,
                                                          const int64_t* __restrict__ roff,
                                                          const int64_t* __restrict__ coff) {
  const int64_t total = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int64_t r = e / cols, c = e % cols;
    double* p = dst + base + roff[r] + coff[c];
    *p = fma(alpha, src[e], *p);
  }
}

struct Shape {
  int nd;
  int64_t len[EIN_MAX_DIMS];
  int64_t str[EIN_MAX_DIMS];
};

// every element of the strided tensor C times beta (MArray.ScaledBy; beta = 0 stores 0)
__global__ __launch_bounds__(256) void tensor_scale_kernel(int64_t total, Shape sh, double beta,
                                                           double* __restrict__ c) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    int64_t rem = e, off = 0;
    for (int d = sh.nd - 1; d >= 0; --d) {
      off += (rem % sh.len[d]) * sh.str[d];
      rem /= sh.len[d];
    }
    c[off] = beta == 0.0 ? 0.0 : beta * c[off];
  }
}

int grid_of(int64_t total) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
}

}  // namespace

int dcontract_internal(const char* expr, double alpha, const double* a, int na,
                       const int64_t* len_a, const int64_t* str_a, const double* b, int nb,
                       const int64_t* len_b, const int64_t* str_b, double beta, double* c,
                       int nc, const int64_t* len_c, const int64_t* str_c) {
  hipStream_t s = stream();
  if (na > EIN_MAX_DIMS || nb > EIN_MAX_DIMS || nc > EIN_MAX_DIMS || na < 0 || nb < 0 || nc < 0) {
    set_error("tensors of at most %d dimensions", EIN_MAX_DIMS);
    return DDLA_ERR_INVALID_ARG;
  }
  // expression "a,b->c" (MArray.Contract.cs:55-63)
  const std::string e(expr);
  const size_t arrow = e.find("->");
  if (arrow == std::string::npos || e.find("->", arrow + 2) != std::string::npos) {
    set_error("Expression must contain exactly one '->' symbol.");
    return DDLA_ERR_INVALID_ARG;
  }
  const std::string lhs = e.substr(0, arrow), symc = trim(e.substr(arrow + 2));
  const size_t comma = lhs.find(',');
  if (comma == std::string::npos || lhs.find(',', comma + 1) != std::string::npos) {
    set_error("Expression must contain exactly two symbols before '->'.");
    return DDLA_ERR_INVALID_ARG;
  }
  const std::string syma = trim(lhs.substr(0, comma)), symb = trim(lhs.substr(comma + 1));
  std::map<char, std::pair<int64_t, int64_t>> ia, ib, ic;
  DDLA_RETURN_IF(diagonal(syma, na, len_a, str_a, ia, "left"));
  DDLA_RETURN_IF(diagonal(symb, nb, len_b, str_b, ib, "right"));
  DDLA_RETURN_IF(diagonal(symc, nc, len_c, str_c, ic, "result"));

  // C = beta C over every element of C (Contract(beta, result, ...) scales first)
  int64_t total_c = 1;
  for (int d = 0; d < nc; ++d) total_c *= len_c[d];
  if (beta != 1.0 && total_c > 0) {
    Shape sh;
    sh.nd = nc;
    for (int d = 0; d < nc; ++d) {
      sh.len[d] = len_c[d];
      sh.str[d] = str_c[d];
    }
    hipLaunchKernelGGL(tensor_scale_kernel, dim3(grid_of(total_c)), dim3(256), 0, s, total_c, sh,
                       beta, c);
    DDLA_HIP_CHECK(hipGetLastError());
  }
  if (alpha == 0.0) return 0;

  // index classes (IndiceUtils.Divide)
  std::vector<Idx> M, N, K, batch;
  auto len_mismatch = [](char ch) {
    set_error("The character '%c' exists in both indices with different lengths.", ch);
    return DDLA_ERR_INVALID_ARG;
  };
  for (auto& [ch, la] : ia) {
    auto jb = ib.find(ch);
    auto jc = ic.find(ch);
    Idx x;
    x.len = la.first;
    x.sa = la.second;
    if (jb != ib.end() && jb->second.first != x.len) return len_mismatch(ch);
    if (jc != ic.end() && jc->second.first != x.len) return len_mismatch(ch);
    if (jb != ib.end()) x.sb = jb->second.second;
    if (jc != ic.end()) x.sc = jc->second.second;
    if (jb != ib.end() && jc != ic.end())
      batch.push_back(x);  // A, B and C
    else if (jb != ib.end())
      K.push_back(x);  // A and B
    else if (jc != ic.end())
      M.push_back(x);  // A and C
    else
      batch.push_back(x);  // A alone: summed
  }
  for (auto& [ch, lb] : ib) {
    if (ia.count(ch)) continue;
    auto jc = ic.find(ch);
    Idx x;
    x.len = lb.first;
    x.sb = lb.second;
    if (jc != ic.end() && jc->second.first != x.len) return len_mismatch(ch);
    if (jc != ic.end()) {
      x.sc = jc->second.second;
      N.push_back(x);  // B and C
    } else {
      batch.push_back(x);  // B alone: summed
    }
  }
  for (auto& [ch, lc] : ic) {
    if (ia.count(ch) || ib.count(ch)) continue;
    Idx x;
    x.len = lc.first;
    x.sc = lc.second;
    batch.push_back(x);  // C alone: broadcast
  }
  for (const auto* v : {&M, &N, &K, &batch})
    for (const Idx& d : *v)
      if (d.len == 0) return 0;  // an empty index: nothing to add

  std::vector<int64_t> rA, rC, cB, cC, kA, kB, bA, bB, bC;
  offsets(M, &Idx::sa, rA);
  offsets(M, &Idx::sc, rC);
  offsets(N, &Idx::sb, cB);
  offsets(N, &Idx::sc, cC);
  offsets(K, &Idx::sa, kA);
  offsets(K, &Idx::sb, kB);
  offsets(batch, &Idx::sa, bA);
  offsets(batch, &Idx::sb, bB);
  offsets(batch, &Idx::sc, bC);
  const int64_t m = (int64_t)rA.size(), n = (int64_t)cB.size(), k = (int64_t)kA.size();

  // scatter vectors on the device, then the three operand buffers
  const size_t nidx = rA.size() + rC.size() + cB.size() + cC.size() + kA.size() + kB.size();
  void* ws = nullptr;
  DDLA_RETURN_IF(workspace(WS_EIN_IDX, nidx * sizeof(int64_t), &ws));
  int64_t* d_idx = static_cast<int64_t*>(ws);
  std::vector<int64_t> host;
  host.reserve(nidx);
  for (const auto* v : {&rA, &rC, &cB, &cC, &kA, &kB}) host.insert(host.end(), v->begin(), v->end());
  DDLA_HIP_CHECK(hipMemcpyAsync(d_idx, host.data(), nidx * sizeof(int64_t),
                                hipMemcpyHostToDevice, s));
  const int64_t* drA = d_idx;
  const int64_t* drC = drA + m;
  const int64_t* dcB = drC + m;
  const int64_t* dcC = dcB + n;
  const int64_t* dkA = dcC + n;
  const int64_t* dkB = dkA + k;
  void *wa = nullptr, *wb = nullptr, *wc = nullptr;
  DDLA_RETURN_IF(workspace(WS_EIN_A, (size_t)(m * k) * sizeof(double), &wa));
  DDLA_RETURN_IF(workspace(WS_EIN_B, (size_t)(k * n) * sizeof(double), &wb));
  DDLA_RETURN_IF(workspace(WS_EIN_C, (size_t)(m * n) * sizeof(double), &wc));
  double* Ap = static_cast<double*>(wa);
  double* Bp = static_cast<double*>(wb);
  double* Cp = static_cast<double*>(wc);
  for (size_t q = 0; q < bA.size(); ++q) {
    hipLaunchKernelGGL(gather_kernel, dim3(grid_of(m * k)), dim3(256), 0, s, m, k, a, bA[q], drA,
                       dkA, Ap);
    hipLaunchKernelGGL(gather_kernel, dim3(grid_of(k * n)), dim3(256), 0, s, k, n, b, bB[q], dkB,
                       dcB, Bp);
    DDLA_HIP_CHECK(hipGetLastError());
    DDLA_RETURN_IF(launch_dgemm(m, n, k, 1.0, Ap, k, true, Bp, n, true, 0.0, Cp, n, 0, s));
    hipLaunchKernelGGL(scatter_add_kernel, dim3(grid_of(m * n)), dim3(256), 0, s, m, n, Cp, alpha,
                       c, bC[q], drC, dcC);
    DDLA_HIP_CHECK(hipGetLastError());
  }
  // the host copy of the scatter vectors must outlive the transfer
  DDLA_HIP_CHECK(hipStreamSynchronize(s));
  return 0;
}

}  // namespace ddla
