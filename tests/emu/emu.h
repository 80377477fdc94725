// CPU stand-in for the HIP built-ins used by the workgroup kernels of the e
// pipeline (k_bsr_subtree, k_radix_subtree), so that their exact integer
// results can be checked on a machine without a GPU: one std::thread per
// lane, std::barrier for __syncthreads, __shared__ arrays as statics (one
// workgroup runs at a time).
#include <barrier>
#include <cstdint>
#include <cstdio>
#include <thread>
#include <vector>
struct Dim {
  unsigned x = 1, y = 1, z = 1;
};
thread_local Dim threadIdx;
Dim blockIdx, blockDim, gridDim;
std::barrier<> *g_bar;
inline void __syncthreads() { g_bar->arrive_and_wait(); }
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__ static
template <class F>
void run_grid(unsigned grid, unsigned block, F f) {
  for (unsigned b = 0; b < grid; b++) {
    blockIdx.x = b;
    blockDim.x = block;
    gridDim.x = grid;
    std::barrier<> bar(block);
    g_bar = &bar;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < block; t++) th.emplace_back([=] {
        threadIdx.x = t;
        f();
      });
    for (auto &x : th) x.join();
  }
}
