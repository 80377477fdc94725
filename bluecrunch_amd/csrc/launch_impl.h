// Launcher templates shared by the launch_*.hip translation units.
#pragma once
#include "launch.h"

namespace bcg {

#define BCG_LAUNCH2(kern, gx, gy, block, stream, ...) \
  hipLaunchKernelGGL(kern, dim3(gx, gy), dim3(block), 0, stream, __VA_ARGS__)

// Packing/unpacking passes are always the outermost pass, a power of two
// (the 3 * 2^k factor sits in the innermost pass): the flag is rejected on
// 3 * 2^k tiles.
template <int L, int C>
int launch_fwd(bool pack, int grid, int gy, hipStream_t s, const PassArgs &a) {
  if constexpr (Len<L>::THREE) {
    if (pack) return -2;
    BCG_LAUNCH2((k_fwd<L, C, false>), grid, gy, (Tile<L, C>::T), s, a);
  } else {
    if (pack && a.lower)
      BCG_LAUNCH2((k_fwd<L, C, true, 0, true>), grid, gy, (Tile<L, C>::T), s, a);
    else if (pack)
      BCG_LAUNCH2((k_fwd<L, C, true>), grid, gy, (Tile<L, C>::T), s, a);
    else
      BCG_LAUNCH2((k_fwd<L, C, false>), grid, gy, (Tile<L, C>::T), s, a);
  }
  return 0;
}

// Unpacking into group sums (a.ybuf) exists for 8- and 16-column tiles.
template <int L, int C>
int launch_inv(bool unpack, int grid, int gy, hipStream_t s, const PassArgs &a) {
  if constexpr (Len<L>::THREE) {
    if (unpack) return -2;
    BCG_LAUNCH2((k_inv<L, C, 0>), grid, gy, (Tile<L, C>::T), s, a);
  } else {
    if (unpack && a.ybuf) {
      if constexpr (C == 8 || C == 16)
        BCG_LAUNCH2((k_inv<L, C, 2>), grid, gy, (Tile<L, C>::T), s, a);
      else
        return -2;
    } else if (unpack) {
      BCG_LAUNCH2((k_inv<L, C, 1>), grid, gy, (Tile<L, C>::T), s, a);
    } else {
      BCG_LAUNCH2((k_inv<L, C, 0>), grid, gy, (Tile<L, C>::T), s, a);
    }
  }
  return 0;
}

// Column counts with L * C <= 8192 elements per tile (32 columns only for the
// middle inverse pass of 256-point tiles, see inv_cols).
template <int L, template <int, int> class F>
struct ByC {
  static fwd_fn pick(int C) {
    switch (C) {
      case 1: return &F<L, 1>::run;
      case 2: return (L * 2 <= 8192) ? &F<L, (L * 2 <= 8192 ? 2 : 1)>::run : nullptr;
      case 4: return (L * 4 <= 8192) ? &F<L, (L * 4 <= 8192 ? 4 : 1)>::run : nullptr;
      case 8: return (L * 8 <= 8192) ? &F<L, (L * 8 <= 8192 ? 8 : 1)>::run : nullptr;
      case 16: return (L * 16 <= 8192) ? &F<L, (L * 16 <= 8192 ? 16 : 1)>::run : nullptr;
      case 32: return (L == 256) ? &F<L, (L == 256 ? 32 : 1)>::run : nullptr;
    }
    return nullptr;
  }
};

template <int L, int C>
struct FwdL {
  static int run(bool f, int gx, int gy, hipStream_t s, const PassArgs &a) { return launch_fwd<L, C>(f, gx, gy, s, a); }
};
template <int L, int C>
struct InvL {
  static int run(bool f, int gx, int gy, hipStream_t s, const PassArgs &a) { return launch_inv<L, C>(f, gx, gy, s, a); }
};

}  // namespace bcg
