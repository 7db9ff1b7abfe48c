// C ABI entry points (include/ddla_native.h).
#include <algorithm>
#include <cstring>
#include <vector>

#include "ddla_internal.hpp"

namespace ddla {
int init_device(int device);
int finalize();
}  // namespace ddla

using namespace ddla;

namespace ddla {

bool valid_trans(int t) {
  return t == DDLA_NO_TRANSPOSE || t == DDLA_TRANSPOSE || t == DDLA_CONJ_NO_TRANSPOSE ||
         t == DDLA_CONJ_TRANSPOSE;
}
// real data: conjugation is the identity (BLIS ignores conj for real types)
bool is_trans(int t) { return t == DDLA_TRANSPOSE || t == DDLA_CONJ_TRANSPOSE; }

// strided host view -> contiguous row-major host buffer
void pack_host(int64_t m, int64_t n, const double* src, int64_t rs, int64_t cs, double* dst) {
  for (int64_t i = 0; i < m; ++i)
    for (int64_t j = 0; j < n; ++j) dst[i * n + j] = src[i * rs + j * cs];
}
void unpack_host(int64_t m, int64_t n, const double* src, double* dst, int64_t rs, int64_t cs) {
  for (int64_t i = 0; i < m; ++i)
    for (int64_t j = 0; j < n; ++j) dst[i * rs + j * cs] = src[i * n + j];
}

DevBuf::~DevBuf() {
  if (p) hipFree(p);
}

int DevBuf::alloc(size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    p = nullptr;
    set_error("hipMalloc of %zu bytes failed", bytes);
    return DDLA_ERR_ALLOC;
  }
  return 0;
}


int upload_view(int64_t m, int64_t n, const double* host, int64_t rs, int64_t cs, DevBuf& buf) {
  DDLA_RETURN_IF(buf.alloc((size_t)m * n * sizeof(double)));
  if (m * n == 0) return 0;
  if (cs == 1 && rs == n) {
    DDLA_HIP_CHECK(hipMemcpy(buf.p, host, (size_t)m * n * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<double> tmp((size_t)m * n);
    pack_host(m, n, host, rs, cs, tmp.data());
    DDLA_HIP_CHECK(hipMemcpy(buf.p, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return 0;
}

int download_view(int64_t m, int64_t n, const DevBuf& buf, double* host, int64_t rs, int64_t cs) {
  if (m * n == 0) return 0;
  DDLA_HIP_CHECK(hipStreamSynchronize(stream()));
  if (cs == 1 && rs == n) {
    DDLA_HIP_CHECK(hipMemcpy(host, buf.p, (size_t)m * n * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<double> tmp((size_t)m * n);
    DDLA_HIP_CHECK(hipMemcpy(tmp.data(), buf.p, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
    unpack_host(m, n, tmp.data(), host, rs, cs);
  }
  return 0;
}

int trsm_rowmajor(int side, int uplo, int transa, int diag, int64_t m, int64_t n, double alpha,
                  const double* a, int64_t lda, double* b, int64_t ldb) {
  hipStream_t s = stream();
  DDLA_RETURN_IF(launch_dscale2d(m, n, alpha, b, ldb, 1, 0, s));
  const bool lower = uplo == DDLA_LOWER, unit = diag == DDLA_UNIT_DIAG, tr = is_trans(transa);
  if (side == DDLA_LEFT) return trsm_left<double>(lower, unit, tr, m, n, a, lda, b, ldb, s);
  // X op(A) = B  <=>  op(A)^T X^T = B^T
  void* ws = nullptr;
  DDLA_RETURN_IF(workspace(WS_MIXED2, (size_t)m * n * sizeof(double), &ws));
  double* bt = static_cast<double*>(ws);
  DDLA_RETURN_IF(launch_dcopy2d(n, m, b, 1, ldb, bt, m, s));
  DDLA_RETURN_IF(trsm_left<double>(lower, unit, !tr, n, m, a, lda, bt, m, s));
  return launch_dcopy2d(m, n, bt, 1, m, b, ldb, s);
}

// device GEMM with arbitrary strides, reduced to the row-major-C kernel; tri (m == n): 0 the
// whole of C, 1 / 2 only its lower / upper triangle (GEMMT), the other left untouched
int gemm_strided(int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
                 const double* a, int64_t rsa, int64_t csa, const double* b, int64_t rsb,
                 int64_t csb, double beta, double* c, int64_t rsc, int64_t csc, int tri) {
  hipStream_t s = stream();
  if (m == 0 || n == 0) return 0;
  // op(A): m x k with strides (ra, ca); op(B): k x n with strides (rb, cb)
  int64_t ra = is_trans(transa) ? csa : rsa, ca = is_trans(transa) ? rsa : csa;
  int64_t rb = is_trans(transb) ? csb : rsb, cb = is_trans(transb) ? rsb : csb;
  const double* A = a;
  const double* B = b;
  double* C = c;
  int64_t ldc;
  bool c_tmp = false;
  if (csc == 1 || (n == 1 && rsc != 1)) {
    ldc = rsc;  // row-major C (a single column only needs the row stride)
  } else if (rsc == 1) {
    // C^T = op(B)^T op(A)^T on the transposed (row-major) view of C
    std::swap(m, n);
    std::swap(A, B);
    int64_t nra = cb, nca = rb, nrb = ca, ncb = ra;
    ra = nra; ca = nca; rb = nrb; cb = ncb;
    ldc = csc;
    tri = tri == 1 ? 2 : (tri == 2 ? 1 : 0);  // the lower triangle of C is the upper of C^T
  } else {
    c_tmp = true;
    ldc = n;
  }
  if (m == 1 && !c_tmp) ldc = std::max<int64_t>(ldc, 1);
  void* tmp = nullptr;
  size_t need = 0;
  const bool a_ok = (ca == 1) || (ra == 1);
  const bool b_ok = (cb == 1) || (rb == 1);
  const size_t a_bytes = a_ok ? 0 : (size_t)m * k * sizeof(double);
  const size_t b_bytes = b_ok ? 0 : (size_t)k * n * sizeof(double);
  const size_t c_bytes = c_tmp ? (size_t)m * n * sizeof(double) : 0;
  need = a_bytes + b_bytes + c_bytes;
  if (need) DDLA_RETURN_IF(workspace(WS_TEMP, need, &tmp));
  char* tp = static_cast<char*>(tmp);
  bool ak, bn;
  int64_t lda, ldb;
  if (ca == 1) { ak = true; lda = ra; }
  else if (ra == 1) { ak = false; lda = ca; }
  else {
    double* pa = reinterpret_cast<double*>(tp);
    tp += a_bytes;
    DDLA_RETURN_IF(launch_dcopy2d(m, k, A, ra, ca, pa, k, s));
    A = pa; ak = true; lda = k;
  }
  if (cb == 1) { bn = true; ldb = rb; }
  else if (rb == 1) { bn = false; ldb = cb; }
  else {
    double* pb = reinterpret_cast<double*>(tp);
    tp += b_bytes;
    DDLA_RETURN_IF(launch_dcopy2d(k, n, B, rb, cb, pb, n, s));
    B = pb; bn = true; ldb = n;
  }
  // single-row / single-column operands: any positive leading dimension works
  if (lda <= 0) lda = 1;
  if (ldb <= 0) ldb = 1;
  if (c_tmp) {
    double* pc = reinterpret_cast<double*>(tp);
    // the copy is written back whole, so a triangle-only product needs the other triangle too
    if (beta != 0.0 || tri != 0) DDLA_RETURN_IF(launch_dcopy2d(m, n, c, rsc, csc, pc, n, s));
    DDLA_RETURN_IF(launch_dgemm(m, n, k, alpha, A, lda, ak, B, ldb, bn, beta, pc, n, tri, s));
    DDLA_RETURN_IF(launch_dscatter2d(m, n, pc, n, c, rsc, csc, s));
    return 0;
  }
  return launch_dgemm(m, n, k, alpha, A, lda, ak, B, ldb, bn, beta, C, ldc, tri, s);
}

}  // namespace ddla

namespace {

__global__ void amax_kernel(int64_t n, const double* __restrict__ x, int64_t inc,
                            int64_t* __restrict__ out) {
  __shared__ double sk[16];
  __shared__ int64_t sr[16];
  double key = -2.0;
  int64_t row = INT64_MAX;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    double v = fabs(x[i * inc]);
    if (v != v) v = -1.0;
    if (v > key) {  // ascending i within a lane: strict > keeps the first index
      key = v;
      row = i;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double k2 = __shfl_xor(key, off);
    const int64_t r2 = __shfl_xor(row, off);
    if (k2 > key || (k2 == key && r2 < row)) {
      key = k2;
      row = r2;
    }
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sk[wave] = key;
    sr[wave] = row;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      if (sk[w] > key || (sk[w] == key && sr[w] < row)) {
        key = sk[w];
        row = sr[w];
      }
    *out = (row == INT64_MAX) ? 0 : row;
  }
}

__global__ void randm_kernel(int64_t m, int64_t n, double* __restrict__ d, int64_t ld,
                             uint64_t seed) {
  const int64_t total = m * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + static_cast<uint64_t>(t + 1) * 0x9E3779B97F4A7C15ull;
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    const int64_t i = t / n, j = t % n;
    d[i * ld + j] = static_cast<double>(z >> 11) * 0x1.0p-52 - 1.0;
  }
}

}  // namespace

#define API_INIT() DDLA_RETURN_IF(ensure_init())

extern "C" {

int ddla_init(int device) { return init_device(device); }
int ddla_finalize(void) { return finalize(); }

int ddla_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return DDLA_ERR_INVALID_ARG;
  std::strncpy(buf, last_error(), len - 1);
  buf[len - 1] = '\0';
  return 0;
}

int ddla_set_stream(void* s) {
  API_INIT();
  ctx().stream = s ? static_cast<hipStream_t>(s) : ctx().own_stream;
  return 0;
}

int ddla_get_stream(void** s) {
  API_INIT();
  DDLA_ARG_CHECK(s != nullptr, "null output pointer");
  *s = ctx().stream;
  return 0;
}

int ddla_sync(void) {
  API_INIT();
  DDLA_HIP_CHECK(hipStreamSynchronize(stream()));
  return 0;
}

int ddla_version(void) { return 100; }

int ddla_set_wait_limit(uint64_t rounds) {
  API_INIT();
  ctx().wait_limit = rounds == 0 ? DEFAULT_WAIT_LIMIT
                                 : static_cast<unsigned>(std::min<uint64_t>(rounds, 0xFFFFFFFFu));
  ctx().wait_limit_set = true;
  return 0;
}

int ddla_profile_enable(int on) {
  API_INIT();
  profile_enable(on != 0);
  return 0;
}

int ddla_profile_reset(void) {
  API_INIT();
  profile_reset();
  return 0;
}

int ddla_profile_only(const char* name) {
  API_INIT();
  profile_only(name);
  return 0;
}

int ddla_profile_read(int idx, char* name, size_t len, int64_t* count, double* total_ms,
                      double* flops, double* bytes) {
  API_INIT();
  return profile_read(idx, name, len, count, total_ms, flops, bytes);
}

int ddla_malloc(void** dptr, size_t bytes) {
  API_INIT();
  DDLA_ARG_CHECK(dptr != nullptr, "null output pointer");
  if (hipMalloc(dptr, bytes ? bytes : 8) != hipSuccess) {
    *dptr = nullptr;
    set_error("hipMalloc of %zu bytes failed", bytes);
    return DDLA_ERR_ALLOC;
  }
  return 0;
}

int ddla_free(void* dptr) {
  API_INIT();
  if (!dptr) return 0;
  DDLA_HIP_CHECK(hipStreamSynchronize(stream()));
  DDLA_HIP_CHECK(hipFree(dptr));
  return 0;
}

int ddla_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  API_INIT();
  if (!bytes) return 0;
  DDLA_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
  DDLA_HIP_CHECK(hipStreamSynchronize(stream()));
  return 0;
}

int ddla_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  API_INIT();
  if (!bytes) return 0;
  DDLA_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
  DDLA_HIP_CHECK(hipStreamSynchronize(stream()));
  return 0;
}

int ddla_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  API_INIT();
  if (!bytes) return 0;
  DDLA_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
  return 0;
}

int ddla_dmat_set(ddla_dim_t m, ddla_dim_t n, const double* host, ddla_inc_t rs, ddla_inc_t cs,
                  double* dev, ddla_inc_t ld) {
  API_INIT();
  DDLA_ARG_CHECK(m >= 0 && n >= 0 && ld >= n, "bad shape m=%lld n=%lld ld=%lld", (long long)m,
                 (long long)n, (long long)ld);
  if (m == 0 || n == 0) return 0;
  std::vector<double> tmp((size_t)m * n);
  pack_host(m, n, host, rs, cs, tmp.data());
  DDLA_HIP_CHECK(hipMemcpy2DAsync(dev, ld * sizeof(double), tmp.data(), n * sizeof(double),
                                  n * sizeof(double), m, hipMemcpyHostToDevice, stream()));
  DDLA_HIP_CHECK(hipStreamSynchronize(stream()));
  return 0;
}

int ddla_dmat_get(ddla_dim_t m, ddla_dim_t n, const double* dev, ddla_inc_t ld, double* host,
                  ddla_inc_t rs, ddla_inc_t cs) {
  API_INIT();
  DDLA_ARG_CHECK(m >= 0 && n >= 0 && ld >= n, "bad shape m=%lld n=%lld ld=%lld", (long long)m,
                 (long long)n, (long long)ld);
  if (m == 0 || n == 0) return 0;
  std::vector<double> tmp((size_t)m * n);
  DDLA_HIP_CHECK(hipMemcpy2DAsync(tmp.data(), n * sizeof(double), dev, ld * sizeof(double),
                                  n * sizeof(double), m, hipMemcpyDeviceToHost, stream()));
  DDLA_HIP_CHECK(hipStreamSynchronize(stream()));
  unpack_host(m, n, tmp.data(), host, rs, cs);
  return 0;
}

int ddla_drandm_dev(ddla_dim_t m, ddla_dim_t n, double* dev, ddla_inc_t ld, uint64_t seed) {
  API_INIT();
  DDLA_ARG_CHECK(m >= 0 && n >= 0 && ld >= n, "bad shape");
  if (m == 0 || n == 0) return 0;
  int64_t g = std::min<int64_t>((m * n + 255) / 256, 65536);
  hipLaunchKernelGGL(randm_kernel, dim3((unsigned)g), dim3(256), 0, stream(), m, n, dev, ld, seed);
  DDLA_HIP_CHECK(hipGetLastError());
  return 0;
}

int ddla_dgemm_dev(int transa, int transb, ddla_dim_t m, ddla_dim_t n, ddla_dim_t k,
                   const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa,
                   const double* b, ddla_inc_t rsb, ddla_inc_t csb, const double* beta, double* c,
                   ddla_inc_t rsc, ddla_inc_t csc) {
  API_INIT();
  DDLA_ARG_CHECK(valid_trans(transa) && valid_trans(transb), "bad trans flag");
  DDLA_ARG_CHECK(m >= 0 && n >= 0 && k >= 0, "negative dimension");
  DDLA_ARG_CHECK(alpha && beta, "null alpha/beta");
  return gemm_strided(transa, transb, m, n, k, *alpha, a, rsa, csa, b, rsb, csb, *beta, c, rsc,
                      csc, 0);
}

int ddla_dgemm(int transa, int transb, ddla_dim_t m, ddla_dim_t n, ddla_dim_t k,
               const double* alpha, const double* a, ddla_inc_t rsa, ddla_inc_t csa, const double* b,
               ddla_inc_t rsb, ddla_inc_t csb, const double* beta, double* c, ddla_inc_t rsc,
               ddla_inc_t csc) {
  API_INIT();
  DDLA_ARG_CHECK(valid_trans(transa) && valid_trans(transb), "bad trans flag");
  DDLA_ARG_CHECK(m >= 0 && n >= 0 && k >= 0, "negative dimension");
  DDLA_ARG_CHECK(alpha && beta, "null alpha/beta");
  if (m == 0 || n == 0) return 0;
  const int64_t am = is_trans(transa) ? k : m, an = is_trans(transa) ? m : k;
  const int64_t bm = is_trans(transb) ? n : k, bn = is_trans(transb) ? k : n;
  DevBuf dA, dB, dC;
  DDLA_RETURN_IF(upload_view(am, an, a, rsa, csa, dA));
  DDLA_RETURN_IF(upload_view(bm, bn, b, rsb, csb, dB));
  if (*beta != 0.0) {
    DDLA_RETURN_IF(upload_view(m, n, c, rsc, csc, dC));
  } else {
    DDLA_RETURN_IF(dC.alloc((size_t)m * n * sizeof(double)));
  }
  DDLA_RETURN_IF(gemm_strided(transa, transb, m, n, k, *alpha, dA.d(), an, 1, dB.d(), bn, 1,
                              *beta, dC.d(), n, 1, 0));
  return download_view(m, n, dC, c, rsc, csc);
}

int ddla_dgetrf_dev(ddla_dim_t n, double* a, ddla_inc_t lda, int32_t* dpiv) {
  API_INIT();
  DDLA_ARG_CHECK(n >= 0 && lda >= n && (n == 0 || (a && dpiv)), "bad getrf arguments");
  if (n == 0) return 0;
  DDLA_ARG_CHECK(n <= INT32_MAX, "n too large");
  int32_t info = 0;
  DDLA_RETURN_IF(dgetrf_internal(n, a, lda, dpiv, &info));
  return info;
}

int ddla_dgetrs_dev(int trans, ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t lda,
                    const int32_t* dpiv, double* b, ddla_inc_t ldb) {
  API_INIT();
  DDLA_ARG_CHECK(valid_trans(trans), "bad trans flag");
  DDLA_ARG_CHECK(n >= 0 && nrhs >= 0 && lda >= n && ldb >= nrhs, "bad getrs arguments");
  if (n == 0 || nrhs == 0) return 0;
  DDLA_ARG_CHECK(a && dpiv && b, "null pointer");
  DDLA_RETURN_IF(check_pivots(n, dpiv, stream()));
  return dgetrs_internal(is_trans(trans), n, nrhs, a, lda, dpiv, b, ldb);
}

int ddla_dgetrf(ddla_dim_t n, double* a, ddla_inc_t rsa, ddla_inc_t csa, int32_t* piv) {
  API_INIT();
  DDLA_ARG_CHECK(n >= 0, "negative n");
  if (n == 0) return 0;
  DevBuf dA, dP;
  DDLA_RETURN_IF(upload_view(n, n, a, rsa, csa, dA));
  DDLA_RETURN_IF(dP.alloc((size_t)n * sizeof(int32_t)));
  int32_t info = 0;
  DDLA_RETURN_IF(dgetrf_internal(n, dA.d(), n, static_cast<int32_t*>(dP.p), &info));
  DDLA_RETURN_IF(download_view(n, n, dA, a, rsa, csa));
  DDLA_HIP_CHECK(hipMemcpy(piv, dP.p, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
  return info;
}

int ddla_dgetrs(int trans, ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t rsa,
                ddla_inc_t csa, const int32_t* piv, double* b, ddla_inc_t rsb, ddla_inc_t csb) {
  API_INIT();
  DDLA_ARG_CHECK(valid_trans(trans) && n >= 0 && nrhs >= 0, "bad getrs arguments");
  if (n == 0 || nrhs == 0) return 0;
  DDLA_ARG_CHECK(a && piv && b, "null pointer");
  for (int64_t i = 0; i < n; ++i)
    DDLA_ARG_CHECK(piv[i] <= 0 || i + piv[i] < n,
                   "the pivot at index %lld swaps outside the %lld rows (SwapRow throws there)",
                   (long long)i, (long long)n);
  DevBuf dA, dB, dP;
  DDLA_RETURN_IF(upload_view(n, n, a, rsa, csa, dA));
  DDLA_RETURN_IF(upload_view(n, nrhs, b, rsb, csb, dB));
  DDLA_RETURN_IF(dP.alloc((size_t)n * sizeof(int32_t)));
  DDLA_HIP_CHECK(hipMemcpy(dP.p, piv, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
  DDLA_RETURN_IF(dgetrs_internal(is_trans(trans), n, nrhs, dA.d(), n,
                                 static_cast<int32_t*>(dP.p), dB.d(), nrhs));
  return download_view(n, nrhs, dB, b, rsb, csb);
}

int ddla_dlaswp_dev(int side, int forward, ddla_dim_t m, ddla_dim_t n, double* x, ddla_inc_t ldx,
                    const int32_t* dpiv, ddla_dim_t npiv) {
  API_INIT();
  DDLA_ARG_CHECK(side == DDLA_LEFT || side == DDLA_RIGHT, "bad side");
  DDLA_ARG_CHECK(m >= 0 && n >= 0 && ldx >= n, "bad shape");
  DDLA_ARG_CHECK(npiv == (side == DDLA_LEFT ? m : n),
                 "the length of pivs must match the permuted dimension (LU.cs:240-247)");
  if (m == 0 || n == 0) return 0;
  DDLA_ARG_CHECK(x && dpiv, "null pointer");
  hipStream_t s = stream();
  DDLA_RETURN_IF(check_pivots(npiv, dpiv, s));
  void* ws = nullptr;
  const size_t perm_bytes = ((size_t)npiv * sizeof(int) + 255) / 256 * 256;
  DDLA_RETURN_IF(workspace(WS_REFINE, perm_bytes + (size_t)m * n * sizeof(double), &ws));
  int* perm = static_cast<int*>(ws);
  double* tmp = reinterpret_cast<double*>(static_cast<char*>(ws) + perm_bytes);
  DDLA_RETURN_IF(launch_piv_to_perm(npiv, dpiv, perm, s));
  if (side == DDLA_LEFT) {
    DDLA_RETURN_IF(launch_gather_rows<double>(m, n, x, ldx, perm, tmp, n, !forward, s));
    DDLA_RETURN_IF(launch_dcopy2d(m, n, tmp, n, 1, x, ldx, s));
  } else {
    // columns: permute the rows of the transposed view
    DDLA_RETURN_IF(launch_dcopy2d(n, m, x, 1, ldx, tmp, m, s));  // tmp = X^T (n x m)
    double* tmp2 = tmp;  // reuse: gather into x^T layout then copy back transposed
    void* ws2 = nullptr;
    DDLA_RETURN_IF(workspace(WS_TEMP, (size_t)m * n * sizeof(double), &ws2));
    double* t2 = static_cast<double*>(ws2);
    DDLA_RETURN_IF(launch_gather_rows<double>(n, m, tmp2, m, perm, t2, m, !forward, s));
    DDLA_RETURN_IF(launch_dcopy2d(m, n, t2, 1, m, x, ldx, s));
  }
  return 0;
}

int ddla_dgetrf_nopiv_dev(ddla_dim_t n, double* a, ddla_inc_t lda) {
  API_INIT();
  DDLA_ARG_CHECK(n >= 0 && lda >= n, "bad getrf arguments");
  if (n == 0) return 0;
  int32_t info = 0;
  DDLA_RETURN_IF(dgetrf_nopiv_internal(n, a, lda, &info));
  return info;
}

int ddla_dgetrs_nopiv_dev(int trans, ddla_dim_t n, ddla_dim_t nrhs, const double* a,
                          ddla_inc_t lda, double* b, ddla_inc_t ldb) {
  API_INIT();
  DDLA_ARG_CHECK(valid_trans(trans), "bad trans flag");
  DDLA_ARG_CHECK(n >= 0 && nrhs >= 0 && lda >= n && ldb >= nrhs, "bad getrs arguments");
  if (n == 0 || nrhs == 0) return 0;
  return dgetrs_nopiv_internal(is_trans(trans), n, nrhs, a, lda, b, ldb);
}

int ddla_dpotrf_dev(int uplo, ddla_dim_t n, double* a, ddla_inc_t lda) {
  API_INIT();
  DDLA_ARG_CHECK(uplo == DDLA_LOWER || uplo == DDLA_UPPER,
                 "Matrix must be either upper or lower triangular.");
  DDLA_ARG_CHECK(n >= 0 && lda >= n, "bad potrf arguments");
  if (n == 0) return 0;
  int32_t info = 0;
  DDLA_RETURN_IF(dpotrf_internal(uplo == DDLA_LOWER, n, a, lda, &info));
  return info;
}

int ddla_dpotrs_dev(int uplo, ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t lda,
                    double* b, ddla_inc_t ldb) {
  API_INIT();
  DDLA_ARG_CHECK(uplo == DDLA_LOWER || uplo == DDLA_UPPER,
                 "Matrix must be either upper or lower triangular.");
  DDLA_ARG_CHECK(n >= 0 && nrhs >= 0 && lda >= n && ldb >= nrhs, "bad potrs arguments");
  if (n == 0 || nrhs == 0) return 0;
  return dpotrs_internal(uplo == DDLA_LOWER, n, nrhs, a, lda, b, ldb);
}

int ddla_dpotrf(int uplo, ddla_dim_t n, double* a, ddla_inc_t rsa, ddla_inc_t csa) {
  API_INIT();
  DDLA_ARG_CHECK(uplo == DDLA_LOWER || uplo == DDLA_UPPER,
                 "Matrix must be either upper or lower triangular.");
  DDLA_ARG_CHECK(n >= 0, "negative n");
  if (n == 0) return 0;
  DevBuf dA;
  DDLA_RETURN_IF(upload_view(n, n, a, rsa, csa, dA));
  int32_t info = 0;
  DDLA_RETURN_IF(dpotrf_internal(uplo == DDLA_LOWER, n, dA.d(), n, &info));
  DDLA_RETURN_IF(download_view(n, n, dA, a, rsa, csa));
  return info;
}

int ddla_dpotrs(int uplo, ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t rsa,
                ddla_inc_t csa, double* b, ddla_inc_t rsb, ddla_inc_t csb) {
  API_INIT();
  DDLA_ARG_CHECK(uplo == DDLA_LOWER || uplo == DDLA_UPPER,
                 "Matrix must be either upper or lower triangular.");
  if (n == 0 || nrhs == 0) return 0;
  DevBuf dA, dB;
  DDLA_RETURN_IF(upload_view(n, n, a, rsa, csa, dA));
  DDLA_RETURN_IF(upload_view(n, nrhs, b, rsb, csb, dB));
  DDLA_RETURN_IF(dpotrs_internal(uplo == DDLA_LOWER, n, nrhs, dA.d(), n, dB.d(), nrhs));
  return download_view(n, nrhs, dB, b, rsb, csb);
}

int ddla_dldlt_dev(int uplo, ddla_dim_t n, double* a, ddla_inc_t lda) {
  API_INIT();
  DDLA_ARG_CHECK(uplo == DDLA_LOWER || uplo == DDLA_UPPER,
                 "Matrix must be either upper or lower triangular.");
  DDLA_ARG_CHECK(n >= 0 && lda >= n, "bad ldlt arguments");
  if (n == 0) return 0;
  DDLA_ARG_CHECK(a != nullptr, "null matrix pointer");
  int32_t info = 0;
  DDLA_RETURN_IF(dldlt_internal(uplo == DDLA_LOWER, n, a, lda, &info));
  return info;
}

int ddla_dldlt_solve_dev(int uplo, ddla_dim_t n, ddla_dim_t nrhs, const double* a,
                         ddla_inc_t lda, double* b, ddla_inc_t ldb) {
  API_INIT();
  DDLA_ARG_CHECK(uplo == DDLA_LOWER || uplo == DDLA_UPPER,
                 "Matrix must be either upper or lower triangular.");
  DDLA_ARG_CHECK(n >= 0 && nrhs >= 0 && lda >= n && ldb >= nrhs, "bad ldlt solve arguments");
  if (n == 0 || nrhs == 0) return 0;
  DDLA_ARG_CHECK(a && b, "null matrix pointer");
  return dldlt_solve_internal(uplo == DDLA_LOWER, n, nrhs, a, lda, b, ldb);
}

int ddla_dldlt(int uplo, ddla_dim_t n, double* a, ddla_inc_t rsa, ddla_inc_t csa) {
  API_INIT();
  DDLA_ARG_CHECK(uplo == DDLA_LOWER || uplo == DDLA_UPPER,
                 "Matrix must be either upper or lower triangular.");
  DDLA_ARG_CHECK(n >= 0, "negative n");
  if (n == 0) return 0;
  DevBuf dA;
  DDLA_RETURN_IF(upload_view(n, n, a, rsa, csa, dA));
  int32_t info = 0;
  DDLA_RETURN_IF(dldlt_internal(uplo == DDLA_LOWER, n, dA.d(), n, &info));
  DDLA_RETURN_IF(download_view(n, n, dA, a, rsa, csa));
  return info;
}

int ddla_dldlt_solve(int uplo, ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t rsa,
                     ddla_inc_t csa, double* b, ddla_inc_t rsb, ddla_inc_t csb) {
  API_INIT();
  DDLA_ARG_CHECK(uplo == DDLA_LOWER || uplo == DDLA_UPPER,
                 "Matrix must be either upper or lower triangular.");
  DDLA_ARG_CHECK(n >= 0 && nrhs >= 0, "negative size");
  if (n == 0 || nrhs == 0) return 0;
  DevBuf dA, dB;
  DDLA_RETURN_IF(upload_view(n, n, a, rsa, csa, dA));
  DDLA_RETURN_IF(upload_view(n, nrhs, b, rsb, csb, dB));
  DDLA_RETURN_IF(dldlt_solve_internal(uplo == DDLA_LOWER, n, nrhs, dA.d(), n, dB.d(), nrhs));
  return download_view(n, nrhs, dB, b, rsb, csb);
}

int ddla_dcontract_dev(const char* expr, double alpha, const double* a, int na,
                       const int64_t* len_a, const int64_t* str_a, const double* b, int nb,
                       const int64_t* len_b, const int64_t* str_b, double beta, double* c,
                       int nc, const int64_t* len_c, const int64_t* str_c) {
  API_INIT();
  DDLA_ARG_CHECK(expr != nullptr, "null expression");
  DDLA_ARG_CHECK((na == 0 || (len_a && str_a)) && (nb == 0 || (len_b && str_b)) &&
                     (nc == 0 || (len_c && str_c)),
                 "null shape arrays");
  DDLA_ARG_CHECK(a && b && c, "null tensor pointer");
  return dcontract_internal(expr, alpha, a, na, len_a, str_a, b, nb, len_b, str_b, beta, c, nc,
                            len_c, str_c);
}

int ddla_qr_t_rows(ddla_dim_t m, ddla_dim_t n) {
  if (m < 0 || n < 0) return 0;
  return qr_t_rows(m, n);
}

int ddla_dgeqrf_dev(ddla_dim_t m, ddla_dim_t n, double* a, ddla_inc_t lda, double* t,
                    ddla_inc_t ldt) {
  API_INIT();
  DDLA_ARG_CHECK(m >= 0 && n >= 0 && lda >= n && ldt >= n, "bad geqrf arguments");
  if (m == 0 || n == 0) return 0;
  DDLA_ARG_CHECK(a && t, "null matrix pointer");
  return dgeqrf_internal(m, n, a, lda, t, ldt);
}

int ddla_dqr_apply_dev(int trans, ddla_dim_t m, ddla_dim_t n, const double* a, ddla_inc_t lda,
                       const double* t, ddla_inc_t ldt, ddla_dim_t nrhs, double* b,
                       ddla_inc_t ldb) {
  API_INIT();
  DDLA_ARG_CHECK(valid_trans(trans), "bad trans");
  DDLA_ARG_CHECK(m >= 0 && n >= 0 && nrhs >= 0 && lda >= n && ldt >= n && ldb >= nrhs,
                 "bad QR apply arguments");
  if (m == 0 || n == 0 || nrhs == 0) return 0;
  DDLA_ARG_CHECK(a && t && b, "null matrix pointer");
  return dqr_apply_internal(is_trans(trans), m, n, a, lda, t, ldt, nrhs, b, ldb);
}

int ddla_dqr_solve_dev(ddla_dim_t m, ddla_dim_t n, ddla_dim_t nrhs, const double* a,
                       ddla_inc_t lda, const double* t, ddla_inc_t ldt, const double* b,
                       ddla_inc_t ldb, double* x, ddla_inc_t ldx) {
  API_INIT();
  DDLA_ARG_CHECK(m >= n, "QR solve needs m >= n (least squares)");
  DDLA_ARG_CHECK(n >= 0 && nrhs >= 0 && lda >= n && ldt >= n && ldb >= nrhs && ldx >= nrhs,
                 "bad QR solve arguments");
  if (n == 0 || nrhs == 0) return 0;
  DDLA_ARG_CHECK(a && t && b && x, "null matrix pointer");
  return dqr_solve_internal(m, n, nrhs, a, lda, t, ldt, b, ldb, x, ldx);
}

int ddla_dgeqrf(ddla_dim_t m, ddla_dim_t n, double* a, ddla_inc_t rsa, ddla_inc_t csa,
                double* t, ddla_inc_t rst, ddla_inc_t cst) {
  API_INIT();
  DDLA_ARG_CHECK(m >= 0 && n >= 0, "negative size");
  if (m == 0 || n == 0) return 0;
  const int64_t tr = qr_t_rows(m, n);
  DevBuf dA, dT;
  DDLA_RETURN_IF(upload_view(m, n, a, rsa, csa, dA));
  DDLA_RETURN_IF(upload_view(tr, n, t, rst, cst, dT));
  DDLA_RETURN_IF(dgeqrf_internal(m, n, dA.d(), n, dT.d(), n));
  DDLA_RETURN_IF(download_view(m, n, dA, a, rsa, csa));
  return download_view(tr, n, dT, t, rst, cst);
}

int ddla_dqr_solve(ddla_dim_t m, ddla_dim_t n, ddla_dim_t nrhs, const double* a, ddla_inc_t rsa,
                   ddla_inc_t csa, const double* t, ddla_inc_t rst, ddla_inc_t cst,
                   const double* b, ddla_inc_t rsb, ddla_inc_t csb, double* x, ddla_inc_t rsx,
                   ddla_inc_t csx) {
  API_INIT();
  DDLA_ARG_CHECK(m >= n, "QR solve needs m >= n (least squares)");
  DDLA_ARG_CHECK(n >= 0 && nrhs >= 0, "negative size");
  if (n == 0 || nrhs == 0) return 0;
  const int64_t tr = qr_t_rows(m, n);
  DevBuf dA, dT, dB, dX;
  DDLA_RETURN_IF(upload_view(m, n, a, rsa, csa, dA));
  DDLA_RETURN_IF(upload_view(tr, n, t, rst, cst, dT));
  DDLA_RETURN_IF(upload_view(m, nrhs, b, rsb, csb, dB));
  DDLA_RETURN_IF(dX.alloc((size_t)n * nrhs * sizeof(double)));
  DDLA_RETURN_IF(dqr_solve_internal(m, n, nrhs, dA.d(), n, dT.d(), n, dB.d(), nrhs, dX.d(), nrhs));
  return download_view(n, nrhs, dX, x, rsx, csx);
}

int ddla_dtrsm_dev(int side, int uplo, int transa, int diag, ddla_dim_t m, ddla_dim_t n,
                   const double* alpha, const double* a, ddla_inc_t lda, double* b,
                   ddla_inc_t ldb) {
  API_INIT();
  DDLA_ARG_CHECK(side == DDLA_LEFT || side == DDLA_RIGHT, "bad side");
  DDLA_ARG_CHECK(uplo == DDLA_LOWER || uplo == DDLA_UPPER, "bad uplo");
  DDLA_ARG_CHECK(valid_trans(transa), "bad trans");
  DDLA_ARG_CHECK(diag == DDLA_UNIT_DIAG || diag == DDLA_NONUNIT_DIAG, "bad diag");
  DDLA_ARG_CHECK(alpha != nullptr, "null alpha");
  const int64_t ka = (side == DDLA_LEFT) ? m : n;
  DDLA_ARG_CHECK(m >= 0 && n >= 0 && lda >= ka && ldb >= n, "bad trsm arguments");
  if (m == 0 || n == 0) return 0;
  DDLA_ARG_CHECK(a && b, "null matrix pointer");
  return trsm_rowmajor(side, uplo, transa, diag, m, n, *alpha, a, lda, b, ldb);
}

int ddla_dsyrk_dev(int uplo, int transa, ddla_dim_t m, ddla_dim_t k, const double* alpha,
                   const double* a, ddla_inc_t lda, const double* beta, double* c, ddla_inc_t ldc) {
  API_INIT();
  DDLA_ARG_CHECK(uplo == DDLA_LOWER || uplo == DDLA_UPPER, "bad uplo");
  DDLA_ARG_CHECK(valid_trans(transa), "bad trans");
  DDLA_ARG_CHECK(alpha && beta, "null alpha/beta");
  DDLA_ARG_CHECK(m >= 0 && k >= 0 && ldc >= m, "bad syrk arguments");
  DDLA_ARG_CHECK(lda >= (is_trans(transa) ? m : k), "lda %lld too small for op(A) of %lld x %lld",
                 (long long)lda, (long long)m, (long long)k);
  if (m == 0) return 0;
  DDLA_ARG_CHECK(c && (k == 0 || a), "null matrix pointer");
  const int tri = (uplo == DDLA_LOWER) ? 1 : 2;
  if (!is_trans(transa))  // C += A A^T, A m x k
    return launch_dgemm(m, m, k, *alpha, a, lda, true, a, lda, false, *beta, c, ldc, tri, stream());
  // C += A^T A, A k x m
  return launch_dgemm(m, m, k, *alpha, a, lda, false, a, lda, true, *beta, c, ldc, tri, stream());
}

int ddla_damaxv_dev(ddla_dim_t n, const double* x, ddla_inc_t incx, ddla_dim_t* index) {
  API_INIT();
  DDLA_ARG_CHECK(index != nullptr, "null index");
  if (n <= 0) {
    *index = 0;
    return 0;
  }
  void* ws = nullptr;
  DDLA_RETURN_IF(workspace(WS_SMALL, 64, &ws));
  hipLaunchKernelGGL(amax_kernel, dim3(1), dim3(1024), 0, stream(), n, x, incx,
                     static_cast<int64_t*>(ws));
  DDLA_HIP_CHECK(hipGetLastError());
  int64_t r = 0;
  DDLA_HIP_CHECK(hipMemcpyAsync(&r, ws, sizeof(int64_t), hipMemcpyDeviceToHost, stream()));
  DDLA_HIP_CHECK(hipStreamSynchronize(stream()));
  *index = r;
  return 0;
}

int ddla_batch_create(ddla_batch_t* plan, int prec, int64_t count, const int32_t* n,
                      const int32_t* kind, void* a, int32_t* piv, int32_t* info, int tune) {
  API_INIT();
  BatchPlan* p = nullptr;
  DDLA_RETURN_IF(batch_create(&p, prec, count, n, kind, a, piv, info, tune));
  *plan = reinterpret_cast<ddla_batch_t>(p);
  return 0;
}

int ddla_batch_execute(ddla_batch_t plan) {
  API_INIT();
  return batch_execute(reinterpret_cast<BatchPlan*>(plan));
}

int ddla_batch_describe(ddla_batch_t plan, char* buf, size_t len) {
  return batch_describe(reinterpret_cast<const BatchPlan*>(plan), buf, len);
}

int ddla_batch_destroy(ddla_batch_t plan) {
  return batch_destroy(reinterpret_cast<BatchPlan*>(plan));
}

}  // extern "C"
