// Wave-level helpers shared by the factorization kernels (CDNA wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

namespace ddla {

// v of lane `lane` (uniform) in every lane
template <typename T>
__device__ __forceinline__ T read_lane(T v, int lane) {
  if constexpr (sizeof(T) == 8) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
    const unsigned hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __builtin_bit_cast(T, (static_cast<unsigned long long>(hi) << 32) | lo);
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
  }
}

// (key, index) of a lane combined with the pair DPP brings from another lane: larger key wins,
// the smaller index among equal keys
template <int CTRL, int ROWMASK, typename T>
__device__ __forceinline__ void dpp_better(T& a, int& k) {
  T a2;
  if constexpr (sizeof(T) == 8) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, a);
    const unsigned long long id = __builtin_bit_cast(unsigned long long, T(-2));
    const int lo = __builtin_amdgcn_update_dpp(static_cast<int>(id), static_cast<int>(b), CTRL,
                                               ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(static_cast<int>(id >> 32),
                                               static_cast<int>(b >> 32), CTRL, ROWMASK, 0xf,
                                               false);
    a2 = __builtin_bit_cast(T, (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                                   static_cast<unsigned>(lo));
  } else {
    a2 = __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, T(-2)),
                                                           __builtin_bit_cast(int, a), CTRL,
                                                           ROWMASK, 0xf, false));
  }
  const int k2 = __builtin_amdgcn_update_dpp(INT_MAX, k, CTRL, ROWMASK, 0xf, false);
  const bool b = a2 > a || (a2 == a && k2 < k);
  a = b ? a2 : a;
  k = b ? k2 : k;
}

// Arg-max of (a, k) over lanes [0, NL) (NL a power of two <= 64; largest a, smallest k among
// equal a): Hillis-Steele over DPP row shifts, then the gfx9 row broadcasts, only as many
// steps as NL needs; every lane ends with lane NL - 1's result.  Lanes with nothing to offer
// pass (-2, any index).
template <int NL, typename T>
__device__ __forceinline__ void lanes_argmax_dpp(T& a, int& k) {
  static_assert(NL >= 1 && NL <= 64 && (NL & (NL - 1)) == 0, "NL: a power of two <= 64");
  if constexpr (NL > 1) dpp_better<0x111, 0xf>(a, k);   // row_shr:1
  if constexpr (NL > 2) dpp_better<0x112, 0xf>(a, k);   // row_shr:2
  if constexpr (NL > 4) dpp_better<0x114, 0xf>(a, k);   // row_shr:4
  if constexpr (NL > 8) dpp_better<0x118, 0xf>(a, k);   // row_shr:8
  if constexpr (NL > 16) dpp_better<0x142, 0xa>(a, k);  // row_bcast:15
  if constexpr (NL > 32) dpp_better<0x143, 0xc>(a, k);  // row_bcast:31
  a = read_lane(a, NL - 1);
  k = __builtin_amdgcn_readlane(k, NL - 1);
}

// The same over all 64 lanes.
template <typename T>
__device__ __forceinline__ void wave_argmax_dpp(T& a, int& k) {
  lanes_argmax_dpp<64>(a, k);
}

// max (MAX) or min of v over lanes [0, NL) by DPP, the result in every lane (uniform)
template <int NL, bool MAX>
__device__ __forceinline__ unsigned lanes_reduce_u32(unsigned v) {
  constexpr unsigned id = MAX ? 0u : 0xFFFFFFFFu;
  auto f = [](unsigned a, unsigned b) { return MAX ? max(a, b) : min(a, b); };
  if constexpr (NL > 1) v = f(v, __builtin_amdgcn_update_dpp(id, v, 0x111, 0xf, 0xf, false));
  if constexpr (NL > 2) v = f(v, __builtin_amdgcn_update_dpp(id, v, 0x112, 0xf, 0xf, false));
  if constexpr (NL > 4) v = f(v, __builtin_amdgcn_update_dpp(id, v, 0x114, 0xf, 0xf, false));
  if constexpr (NL > 8) v = f(v, __builtin_amdgcn_update_dpp(id, v, 0x118, 0xf, 0xf, false));
  if constexpr (NL > 16) v = f(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xa, 0xf, false));
  if constexpr (NL > 32) v = f(v, __builtin_amdgcn_update_dpp(id, v, 0x143, 0xc, 0xf, false));
  return __builtin_amdgcn_readlane(v, NL - 1);
}

// key of a column entry: 0 for rows that are not candidates, 1 for NaN, bits(|x|) + 2
__device__ __forceinline__ unsigned arg_key(float x, bool cand) {
  const unsigned b = __builtin_bit_cast(unsigned, x) & 0x7FFFFFFFu;
  return !cand ? 0u : (b > 0x7F800000u ? 1u : b + 2u);
}

// The winner over lanes [0, NL): the largest key, among equal keys the smallest tie value
// (label); returns the winning lane (-1 when every key is 0) and the key.
template <int NL>
__device__ __forceinline__ int lanes_winner(unsigned key, unsigned tie, int lane, unsigned& kmax) {
  kmax = lanes_reduce_u32<NL, true>(key);
  if (kmax == 0u) return -1;
  const unsigned long long bm = __ballot(lane < NL && key == kmax);
  if (__popcll(bm) == 1) return __ffsll(bm) - 1;
  const unsigned t = lanes_reduce_u32<NL, false>(lane < NL && key == kmax ? tie : 0xFFFFFFFFu);
  return __ffsll(__ballot(lane < NL && key == kmax && tie == t)) - 1;
}

__device__ __forceinline__ unsigned long long arg_key64(double x, bool cand) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x) & 0x7FFFFFFFFFFFFFFFull;
  return !cand ? 0ull : (b > 0x7FF0000000000000ull ? 1ull : b + 2ull);
}

// The winner over lanes [0, NL) of 64-bit keys: the largest key (its high half by one DPP max,
// the low half among the lanes that share it by a second), the smallest tie value among equal
// keys; returns the lane (-1 when every key is 0) and the key.
template <int NL>
__device__ __forceinline__ int lanes_winner64(unsigned long long key, unsigned tie, int lane,
                                              unsigned long long& kmax) {
  const unsigned hi = static_cast<unsigned>(key >> 32), lo = static_cast<unsigned>(key);
  const unsigned mh = lanes_reduce_u32<NL, true>(lane < NL ? hi : 0u);
  const unsigned ml = lanes_reduce_u32<NL, true>(lane < NL && hi == mh ? lo : 0u);
  kmax = (static_cast<unsigned long long>(mh) << 32) | ml;
  if (kmax == 0ull) return -1;
  const unsigned long long bm = __ballot(lane < NL && key == kmax);
  if (__popcll(bm) == 1) return __ffsll(bm) - 1;
  const unsigned t = lanes_reduce_u32<NL, false>(lane < NL && key == kmax ? tie : 0xFFFFFFFFu);
  return __ffsll(__ballot(lane < NL && key == kmax && tie == t)) - 1;
}

}  // namespace ddla
