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

}  // namespace ddla
