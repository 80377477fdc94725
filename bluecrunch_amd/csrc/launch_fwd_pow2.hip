// Forward pass kernels, power-of-two tiles.
#include "launch_impl.h"

namespace bcg {
fwd_fn pick_fwd_pow2(int L, int C) {
  switch (L) {
    case 64: return ByC<64, FwdL>::pick(C);
    case 128: return ByC<128, FwdL>::pick(C);
    case 256: return ByC<256, FwdL>::pick(C);
    case 512: return ByC<512, FwdL>::pick(C);
    case 1024: return ByC<1024, FwdL>::pick(C);
    case 2048: return ByC<2048, FwdL>::pick(C);
    case 4096: return ByC<4096, FwdL>::pick(C);
    case 8192: return ByC<8192, FwdL>::pick(C);
  }
  return nullptr;
}
// Loads this unit's code object now; the runtime otherwise loads it at the
// first launch of one of its kernels, inside a caller's first product.
int preload_fwd_pow2() {
  hipFuncAttributes fa;
  return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&(k_fwd<256, 16, false>))) == hipSuccess ? 0 : -1;
}

}  // namespace bcg
