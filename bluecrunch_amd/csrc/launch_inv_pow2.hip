// Inverse pass kernels, power-of-two tiles.
#include "launch_impl.h"

namespace bcg {
fwd_fn pick_inv_pow2(int L, int C) {
  switch (L) {
    case 64: return ByC<64, InvL>::pick(C);
    case 128: return ByC<128, InvL>::pick(C);
    case 256: return ByC<256, InvL>::pick(C);
    case 512: return ByC<512, InvL>::pick(C);
    case 1024: return ByC<1024, InvL>::pick(C);
    case 2048: return ByC<2048, InvL>::pick(C);
    case 4096: return ByC<4096, InvL>::pick(C);
    case 8192: return ByC<8192, InvL>::pick(C);
  }
  return nullptr;
}
// Loads this unit's code object now; the runtime otherwise loads it at the
// first launch of one of its kernels, inside a caller's first product.
int preload_inv_pow2() {
  hipFuncAttributes fa;
  return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&(k_inv<256, 16, 0>))) == hipSuccess ? 0 : -1;
}

}  // namespace bcg
