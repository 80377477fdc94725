// One-workgroup product kernels (persistent, one length class per launch).
#include "launch_impl.h"

namespace bcg {

template <int L>
int launch_mul1(int grid, hipStream_t s, const ClassArgs &ca) {
  hipLaunchKernelGGL((k_mul1<L>), dim3(grid), dim3(Tile<L, 1>::T), 0, s, ca);
  return 0;
}

mul1_fn pick_mul1(int L) {
  switch (L) {
    case 64: return &launch_mul1<64>;
    case 128: return &launch_mul1<128>;
    case 256: return &launch_mul1<256>;
    case 512: return &launch_mul1<512>;
    case 1024: return &launch_mul1<1024>;
    case 2048: return &launch_mul1<2048>;
    case 4096: return &launch_mul1<4096>;
    case 8192: return &launch_mul1<8192>;
    case 96: return &launch_mul1<96>;
    case 192: return &launch_mul1<192>;
    case 384: return &launch_mul1<384>;
    case 768: return &launch_mul1<768>;
    case 1536: return &launch_mul1<1536>;
    case 3072: return &launch_mul1<3072>;
    case 6144: return &launch_mul1<6144>;
  }
  return nullptr;
}

// Loads this unit's code object now; the runtime otherwise loads it at the
// first launch of one of its kernels, inside a caller's first product.
int preload_mul1() {
  hipFuncAttributes fa;
  return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&k_mul1<1024>)) == hipSuccess ? 0 : -1;
}

}  // namespace bcg
