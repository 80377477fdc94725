// Inverse pass kernels, 3 * 2^k tiles.
#include "launch_impl.h"

namespace bcg {
fwd_fn pick_inv_three(int L, int C) {
  switch (L) {
    case 96: return ByC<96, InvL>::pick(C);
    case 192: return ByC<192, InvL>::pick(C);
    case 384: return ByC<384, InvL>::pick(C);
    case 768: return ByC<768, InvL>::pick(C);
    case 1536: return ByC<1536, InvL>::pick(C);
    case 3072: return ByC<3072, InvL>::pick(C);
    case 6144: return ByC<6144, InvL>::pick(C);
  }
  return nullptr;
}
// Loads this unit's code object now; the runtime otherwise loads it at the
// first launch of one of its kernels, inside a caller's first product.
int preload_inv_three() {
  hipFuncAttributes fa;
  return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&(k_inv<192, 16, 0>))) == hipSuccess ? 0 : -1;
}

}  // namespace bcg
