// Middle passes of chunked three-pass products (the second forward pass, the
// middle pass and the second inverse pass run one group of blocks at a time).
// These are the plain pass kernels with ordinary, cache-allocating accesses
// where the next launch of the same chunk reads the points back: the forward
// pass stores them, the middle pass loads and stores them, the inverse pass
// loads them, so a chunk's points stay in the memory-side cache between the
// three launches. Instantiated for the pass splits of the long transforms.
#include "launch_impl.h"

namespace bcg {

namespace {

template <int L, int C>
int fwd_keep(bool pack, int grid, int gy, hipStream_t s, const PassArgs &a) {
  if (pack) return -2;
  BCG_LAUNCH2((k_fwd<L, C, false, 2>), grid, gy, (Tile<L, C>::T), s, a);
  return 0;
}

template <int L, int C>
int inv_keep(bool unpack, int grid, int gy, hipStream_t s, const PassArgs &a) {
  if (unpack) return -2;
  BCG_LAUNCH2((k_inv<L, C, 0, 1>), grid, gy, (Tile<L, C>::T), s, a);
  return 0;
}

template <int L>
int mid_keep(bool sq, bool pack, int grid, int gy, hipStream_t s, const PassArgs &a) {
  constexpr int T = Tile<L, 1>::T;
  constexpr int BPW = T < 64 ? 64 / T : 1;
  if (pack) return -2;
  if (BPW > 1 && grid % BPW == 0) {
    if (sq)
      BCG_LAUNCH2((k_mid<L, true, false, false, BPW, 3>), grid / BPW, gy, T * BPW, s, a);
    else
      BCG_LAUNCH2((k_mid<L, false, false, false, BPW, 3>), grid / BPW, gy, T * BPW, s, a);
  } else {
    if (sq)
      BCG_LAUNCH2((k_mid<L, true, false, false, 1, 3>), grid, gy, T, s, a);
    else
      BCG_LAUNCH2((k_mid<L, false, false, false, 1, 3>), grid, gy, T, s, a);
  }
  return 0;
}

}  // namespace

fwd_fn pick_fwd_keep(int L, int C) {
  if (L == 256 && C == 16) return &fwd_keep<256, 16>;
  if (L == 128 && C == 16) return &fwd_keep<128, 16>;
  if (L == 512 && C == 8) return &fwd_keep<512, 8>;
  if (L == 1024 && C == 8) return &fwd_keep<1024, 8>;
  return nullptr;
}

fwd_fn pick_inv_keep(int L, int C) {
  if (L == 256 && C == 32) return &inv_keep<256, 32>;
  if (L == 256 && C == 16) return &inv_keep<256, 16>;
  if (L == 128 && C == 16) return &inv_keep<128, 16>;
  if (L == 512 && C == 8) return &inv_keep<512, 8>;
  if (L == 1024 && C == 8) return &inv_keep<1024, 8>;
  return nullptr;
}

mid_fn pick_mid_keep(int L) {
  switch (L) {
    case 192: return &mid_keep<192>;
    case 1024: return &mid_keep<1024>;
  }
  return nullptr;
}

// Loads this unit's code object now; the runtime otherwise loads it at the
// first launch of one of its kernels, inside a caller's first product.
int preload_keep() {
  hipFuncAttributes fa;
  return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&(k_mid<1024, false, false, false, 1, 3>))) ==
                 hipSuccess
             ? 0
             : -1;
}

}  // namespace bcg
