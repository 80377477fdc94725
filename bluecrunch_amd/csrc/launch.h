// Host launchers of the templated transform kernels. Each family is
// instantiated in its own translation unit (launch_*.hip) so the kernels of a
// build compile in parallel; the engine only sees these function pointers.
#pragma once
#include "kernels.h"

namespace bcg {

// fwd/inv: (pack-or-unpack, grid.x, grid.y, stream, args) -> 0, or nonzero
// when the flag is not supported for that tile.
typedef int (*fwd_fn)(bool, int, int, hipStream_t, const PassArgs &);
// mid: (square, pack, grid.x, grid.y, stream, args)
typedef int (*mid_fn)(bool, bool, int, int, hipStream_t, const PassArgs &);
typedef int (*mul1_fn)(int, hipStream_t, const ClassArgs &);

// nullptr when no kernel exists for the (L, C) tile.
fwd_fn pick_fwd_pow2(int L, int C);
fwd_fn pick_fwd_three(int L, int C);
fwd_fn pick_inv_pow2(int L, int C);
fwd_fn pick_inv_three(int L, int C);
mid_fn pick_mid(int L);
mul1_fn pick_mul1(int L);
// Cache-keeping variants for the chunked middle passes of three-pass
// products (forward pass with ordinary stores, middle pass with ordinary
// loads and stores, inverse pass with ordinary loads); nullptr when that
// tile has none.
fwd_fn pick_fwd_keep(int L, int C);
fwd_fn pick_inv_keep(int L, int C);
mid_fn pick_mid_keep(int L);

// Code-object preloads, one per launch_*.hip unit (called at initialisation).
int preload_mul1();
int preload_fwd_pow2();
int preload_fwd_three();
int preload_inv_pow2();
int preload_inv_three();
int preload_mid();
int preload_keep();

inline fwd_fn pick_fwd(int L, int C) { return (L % 3 == 0) ? pick_fwd_three(L, C) : pick_fwd_pow2(L, C); }
inline fwd_fn pick_inv(int L, int C) { return (L % 3 == 0) ? pick_inv_three(L, C) : pick_inv_pow2(L, C); }

}  // namespace bcg
