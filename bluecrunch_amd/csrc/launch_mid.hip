// Middle pass kernels (last forward pass + pointwise product + first inverse
// pass, or a whole one-pass product), every length class.
#include "launch_impl.h"

namespace bcg {

// grid = number of contiguous blocks of L points (1 for a one-pass product).
// Middle passes with blocks of fewer than 64 threads put BPW = 64 / T blocks
// in one workgroup (one full wave) when the block count allows it.
template <int L>
int launch_mid(bool sq, bool pack, int grid, int gy, hipStream_t s, const PassArgs &a) {
  constexpr int T = Tile<L, 1>::T;
  constexpr int BPW = T < 64 ? 64 / T : 1;
  if (pack) {
    if (sq)
      BCG_LAUNCH2((k_mid<L, true, true, true>), grid, gy, T, s, a);
    else
      BCG_LAUNCH2((k_mid<L, false, true, true>), grid, gy, T, s, a);
  } else if (BPW > 1 && grid % BPW == 0) {
    if (sq)
      BCG_LAUNCH2((k_mid<L, true, false, false, BPW>), grid / BPW, gy, T * BPW, s, a);
    else
      BCG_LAUNCH2((k_mid<L, false, false, false, BPW>), grid / BPW, gy, T * BPW, s, a);
  } else {
    if (sq)
      BCG_LAUNCH2((k_mid<L, true, false, false>), grid, gy, T, s, a);
    else
      BCG_LAUNCH2((k_mid<L, false, false, false>), grid, gy, T, s, a);
  }
  return 0;
}

mid_fn pick_mid(int L) {
  switch (L) {
    case 64: return &launch_mid<64>;
    case 128: return &launch_mid<128>;
    case 256: return &launch_mid<256>;
    case 512: return &launch_mid<512>;
    case 1024: return &launch_mid<1024>;
    case 2048: return &launch_mid<2048>;
    case 4096: return &launch_mid<4096>;
    case 8192: return &launch_mid<8192>;
    case 96: return &launch_mid<96>;
    case 192: return &launch_mid<192>;
    case 384: return &launch_mid<384>;
    case 768: return &launch_mid<768>;
    case 1536: return &launch_mid<1536>;
    case 3072: return &launch_mid<3072>;
    case 6144: return &launch_mid<6144>;
  }
  return nullptr;
}

// Loads this unit's code object now; the runtime otherwise loads it at the
// first launch of one of its kernels, inside a caller's first product.
int preload_mid() {
  hipFuncAttributes fa;
  return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&(k_mid<1024, false, false, false>))) == hipSuccess ? 0 : -1;
}

}  // namespace bcg
