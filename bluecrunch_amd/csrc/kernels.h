// Kernels of the FFT multiply: transform passes, the fused middle pass, the
// rounding/unpack pass and the carry resolve. See fft_device.h for the
// transform structure.
#pragma once
#include "fft_device.h"

namespace bcg {

// One product of a batch.
struct MulDesc {
  const uint32_t *a;
  const uint32_t *b;
  uint32_t *c;
  const uint32_t *d;  // optional addend: c = a*b + d (low cl words)
  unsigned long long al, bl, cl, dl;
  int bits;
  int cbits;
  int gk = 0;  // batched group mode: words per group sum of this product's digits (0: none)
};

struct PassArgs {
  cd *data;                 // complex array of the pass (operand A / product)
  const cd *data_b;         // operand B spectrum array (middle pass)
  const uint32_t *wa;       // digits of A come from these words (pack passes)
  const uint32_t *wb;
  unsigned long long al, bl;
  long long *coef;          // unpack output: 2N rounded coefficients
  unsigned long long N;     // transform length
  unsigned long long M;     // pass column count (stride between rows)
  unsigned long long S;     // pass span L * M
  unsigned long long tw_mul;// 4N / S: converts e^{2 pi i / S} exponents to Omega exponents
  const cd *tl;             // e^{2 pi i t / L}, t < L/2
  Roots roots;
  int bits;                 // digit size
  int ntiles;
  unsigned int *margin;     // float bits of the worst |x - round(x)| (atomic max)
  int margin_mask;          // single products: slot (block & mask) * MARGIN_STRIDE; batched: 0
  double inv_n;
  int fast_round;           // every coefficient < 2^51: round with the 1.5 * 2^52 addition
  const MulDesc *desc;      // batched launch: product blockIdx.y, else nullptr
  int operand;              // batched first pass: 0 reads desc.a, 1 reads desc.b
  int coef_T;               // last pass: coefficient j at coef[(j % M) * 2L + j / M]
  uint32_t *ybuf;           // last pass, group mode: digits [C g, C g + C) summed into ybuf[g K .. g K + K)
  int ygK;                  // words per group sum (a multiple of 4: 16-byte stores)
  int pair;                 // forward passes of A and B in one launch: grid.y = 1 runs B (data_b, wb, bl)
  int pcount;               // batched forward passes of both operands: grid.y = 2 pcount, rows >= pcount run B
  int lower;                // packing pass of single products: every operand leaves the upper half empty
  unsigned long long ystride;  // batched group mode: group-sum words per product in ybuf
};

// Batched launches (desc != nullptr): product p = blockIdx.y owns the arrays
// data + p*N, data_b + p*N, coef + p*2N and margin slot p.
__device__ __forceinline__ void bind_product(PassArgs &a) {
  if (!a.desc) {
    if (a.pair && blockIdx.y == 1) {
      a.data = const_cast<cd *>(a.data_b);
      a.wa = a.wb;
      a.al = a.bl;
    }
    return;
  }
  unsigned long long p = blockIdx.y;
  if (a.pcount) {  // both operands in one launch: the upper half of grid.y runs B into data_b
    const bool opb = p >= (unsigned long long)a.pcount;
    if (opb) {
      p -= (unsigned long long)a.pcount;
      a.data = const_cast<cd *>(a.data_b);
    }
    a.operand = opb ? 1 : 0;
    a.data_b = nullptr;
  }
  const MulDesc &d = a.desc[p];
  a.wa = a.operand ? d.b : d.a;
  a.al = a.operand ? d.bl : d.al;
  a.wb = d.b;
  a.bl = d.bl;
  a.bits = d.bits;
  if (a.data) a.data += p * a.N;
  if (a.data_b) a.data_b += p * a.N;
  if (a.coef) a.coef += p * 2 * a.N;
  if (a.ybuf) a.ybuf += p * a.ystride;
  if (a.margin) a.margin += p;
}

// Tile order that keeps adjacent column tiles on one XCD at the same time:
// workgroup b runs on XCD b % 8, so tile = (b % 8) * (ntiles / 8) + b / 8.
__device__ __forceinline__ int tile_of_block(int ntiles) {
  int b = blockIdx.x;
  if (ntiles >= 64 && (ntiles & 7) == 0) return (b & 7) * (ntiles >> 3) + (b >> 3);
  return b;
}

// b-bit digit j of a little-endian word array (0 beyond the last word).
__device__ __forceinline__ double digit_at(const uint32_t *__restrict__ w, unsigned long long wl, unsigned long long j,
                                           int bits) {
  unsigned long long s = j * (unsigned long long)bits;
  unsigned long long wi = s >> 5;
  if (wi >= wl) return 0.0;
  unsigned sh = (unsigned)(s & 31);
  unsigned long long v = w[wi];
  if (wi + 1 < wl) v |= (unsigned long long)w[wi + 1] << 32;
  return (double)((v >> sh) & ((1ull << bits) - 1ull));
}

// Digit pair (a_j, a_{j+N}) of right-angle input point j.
__device__ __forceinline__ cd packed_digits(const PassArgs &a, const uint32_t *__restrict__ w,
                                            unsigned long long wl, unsigned long long j) {
  return make_double2(digit_at(w, wl, j, a.bits), digit_at(w, wl, j + a.N, a.bits));
}

// Right-angle input point j: (a_j + i a_{j+N}) * Omega^j.
__device__ __forceinline__ cd packed_point(const PassArgs &a, const uint32_t *__restrict__ w, unsigned long long wl,
                                           unsigned long long j) {
  return cmul(packed_digits(a, w, wl, j), omega_pow(a.roots, j));
}

__device__ __forceinline__ void margin_update(unsigned int *m, double d) {
  float f = (float)d;
  // non-negative floats order like their bit patterns
  atomicMax(m, __float_as_uint(f));
}

// Nearest integer of x clamped at 0, and the rounding distance into worst.
// Equals fft/convert.c's (uint64_t)(x + 0.5) whenever |x - round x| < 1/2
// (the multiply re-runs with smaller digits from 1/4 on). fast: |x| < 2^51,
// where adding 1.5 * 2^52 leaves round-to-nearest(x) in the low mantissa bits.
__device__ __forceinline__ long long round_coef(double x, bool fast, double &worst) {
  if (fast) {
    // no clamp at 0: a negative result means an error of at least 1/2 in a
    // non-negative coefficient, outside the regime the margin check trusts
    const double magic = 6755399441055744.0;  // 1.5 * 2^52
    const double y = x + magic;
    worst = fmax(worst, fabs(x - (y - magic)));
    return __double_as_longlong(y) - __double_as_longlong(magic);
  }
  worst = fmax(worst, fabs(x - rint(x)));
  const long long v = (long long)(x + 0.5);
  return v > 0 ? v : 0;
}

// Round one output point j = row * M + col, already multiplied by
// Omega^{-j} (f = w_j Omega^{-j}, unscaled), into coefficients c_j (real part)
// and c_{j+N} (imaginary part); returns the rounding margin. L is the pass
// length of the transposed layout (coef_T only).
__device__ __forceinline__ double round_store(const PassArgs &a, cd f, unsigned long long row,
                                              unsigned long long col, unsigned long long L) {
  double m = 0.0;
  const long long cr = round_coef(f.x * a.inv_n, a.fast_round, m);
  const long long ci = round_coef(f.y * a.inv_n, a.fast_round, m);
  if (a.coef_T) {
    // column-major [col][2L] (real part rows, then imaginary)
    a.coef[col * 2 * L + row] = cr;
    a.coef[col * 2 * L + L + row] = ci;
  } else {
    const unsigned long long j = row * a.M + col;
    a.coef[j] = cr;
    a.coef[j + a.N] = ci;
  }
  return m;
}

// Exact sum of C coefficients v[i] 2^{b i + pos0} (pos0 < 32) as K words
// (K >= the words the sum spans) into out[0..K) (LDS or global). pos0 and b
// are the same across the wave (callers derive them from the tile, never
// from the lane), so the shifts and the word emissions are scalar decisions.
template <int C>
__device__ __forceinline__ void group_sum_words(const unsigned long long (&v)[C], int pos0, int b, int K,
                                                uint32_t *out) {
  // 128-bit accumulator w3:w2:w1:w0; bit 0 is bit 32 k of the sum. Each term
  // v 2^pos (v < 2^51, pos < 32) spans three words and enters through one
  // add-with-carry chain.
  unsigned w0 = 0, w1 = 0, w2 = 0, w3 = 0;
  int pos = pos0, k = 0;
#pragma unroll
  for (int i = 0; i < C; i++) {
    if (pos >= 32) {
      out[k++] = w0;
      w0 = w1;
      w1 = w2;
      w2 = w3;
      w3 = 0;
      pos -= 32;
    }
    const unsigned long long lo = v[i] << pos;
    // v >> (64 - pos), 0 for pos = 0, without a select: (v >> 1) >> (63 - pos)
    const unsigned hi = (unsigned)((unsigned)(v[i] >> 33) >> (31 - pos));
    unsigned c0, c1, c2, c3;
    w0 = __builtin_addc(w0, (unsigned)lo, 0u, &c0);
    w1 = __builtin_addc(w1, (unsigned)(lo >> 32), c0, &c1);
    w2 = __builtin_addc(w2, hi, c1, &c2);
    w3 = __builtin_addc(w3, 0u, c2, &c3);
    pos += b;
  }
  for (; k < K; k++) {
    out[k] = w0;
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = 0;
  }
}

// Bit offset (j0 b) & 31 of a group's first digit j0 = t M + m0 (+ N): the
// same for every row t of a tile, since 32 | M b in the passes that sum groups.
__device__ __forceinline__ int group_pos0(const PassArgs &a, unsigned long long m0, int half) {
  return (int)(((m0 + (half ? a.N : 0ull)) * (unsigned long long)a.bits) & 31ull);
}

// Group sums of the last inverse pass (group mode). After the final stage the
// tile slot (row t, column c) holds the rounded coefficient pair (c_j, c_{j+N})
// of j = t M + m0 + c. For each row, the C real parts (digits j0 + [0, C),
// j0 = t M + m0) and the C imaginary parts (j0 + N + [0, C)) each give one
// exact sum sum_i c_i 2^{b i}, written shifted by (j0 b mod 32) as ygK
// little-endian words (ygK a multiple of 4, zero padded) at ybuf[(j0 / C)
// ygK ...]: the words from word floor(j0 b / 32) up. One thread per row; the
// words are assembled in the row's own LDS slots and leave as 16-byte stores.
template <int L, int C>
__device__ __forceinline__ void write_group_sums(const PassArgs &a, cd *lds, unsigned long long m0, int tid) {
  using TC = Tile<L, C>;
  const int b = a.bits;
  const int K = a.ygK;
  for (int t = tid; t < L; t += TC::T) {
    const int s0 = slot<C>(t, 0);  // the row's C slots are contiguous
    unsigned long long re[C], im[C];
#pragma unroll
    for (int i = 0; i < C; i++) {
      const longlong2 v = reinterpret_cast<const longlong2 *>(lds)[s0 + i];
      re[i] = (unsigned long long)v.x;
      im[i] = (unsigned long long)v.y;
    }
    uint32_t *lw = reinterpret_cast<uint32_t *>(lds + s0);
#pragma unroll
    for (int half = 0; half < 2; half++) {
      const unsigned long long j0 = (unsigned long long)t * a.M + m0 + (half ? a.N : 0ull);
      group_sum_words<C>(half ? im : re, group_pos0(a, m0, half), b, K, lw + half * K);
      uint4 *dst = reinterpret_cast<uint4 *>(a.ybuf + (j0 / C) * (unsigned long long)K);
      for (int q = 0; q < K / 4; q++)
        dst[q] = make_uint4(lw[half * K + 4 * q], lw[half * K + 4 * q + 1], lw[half * K + 4 * q + 2],
                            lw[half * K + 4 * q + 3]);
    }
  }
}

// The same for one half (0: real parts, digits j0 + [0, C); 1: imaginary
// parts, j0 + N + [0, C)) from 8-byte coefficient slots.
template <int L, int C>
__device__ __forceinline__ void write_group_half(const PassArgs &a, long long *lc, unsigned long long m0, int tid,
                                                 int half) {
  using TC = Tile<L, C>;
  const int b = a.bits;
  const int K = a.ygK;
  for (int t = tid; t < L; t += TC::T) {
    const int s0 = slot<C>(t, 0);  // the row's C slots are contiguous
    unsigned long long v[C];
#pragma unroll
    for (int i = 0; i < C; i++) v[i] = (unsigned long long)lc[s0 + i];
    uint32_t *lw = reinterpret_cast<uint32_t *>(lc + s0);
    const unsigned long long j0 = (unsigned long long)t * a.M + m0 + (half ? a.N : 0ull);
    group_sum_words<C>(v, group_pos0(a, m0, half), b, K, lw);
    uint4 *dst = reinterpret_cast<uint4 *>(a.ybuf + (j0 / C) * (unsigned long long)K);
    for (int q = 0; q < K / 4; q++) dst[q] = make_uint4(lw[4 * q], lw[4 * q + 1], lw[4 * q + 2], lw[4 * q + 3]);
  }
}

// Whole-point form for single-pass plans (M = 1): j = row.
__device__ __forceinline__ double unpack_point(const PassArgs &a, cd v, unsigned long long j) {
  return round_store(a, cmulc(v, omega_pow(a.roots, j)), j, 0, 0);
}

// Margin slots: a single product spreads its atomics over MARGIN_SLOTS
// addresses MARGIN_STRIDE words apart (one 128-byte line each), so tens of
// thousands of workgroups do not serialise on one address; the host takes the
// maximum over the slots.
constexpr int MARGIN_SLOTS = 64;
constexpr int MARGIN_STRIDE = 32;

// Wave-level max then one atomic per wave, into slot (blockIdx.x & mask).
__device__ __forceinline__ void margin_reduce(unsigned int *m, double d, int mask = 0) {
  float f = (float)d;
  for (int off = 32; off > 0; off >>= 1) f = fmaxf(f, __shfl_xor(f, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(m + (blockIdx.x & mask) * MARGIN_STRIDE, __float_as_uint(f));
}

// ------------------------------------------------------------ forward pass

// Length-L DIF down C adjacent columns of one span; stage 0 loads straight from
// global (or from words when PACK), the rest run in LDS; the store multiplies
// row r (frequency k) by the pass twiddle e^{2 pi i m k / S}.
// Tiles of two radix-16 stages (L = 256): stage 1 runs in registers too, and
// the transpose between the stages goes through LDS one half (real parts,
// then imaginary parts) at a time, so a tile needs half the LDS and twice as
// many waves fit on a CU.
template <int L, int C>
struct SplitTile {
  static constexpr bool ON = !Tile<L, C>::THREE && Tile<L, C>::NST2 == 2 && Tile<L, C>::FIRSTB == 4;
  static constexpr int BYTES = Tile<L, C>::LDS_ELEMS * (ON ? 8 : 16);
};

// Rows r0 + 16 t (t < 16) of column c in x become rows 16 r0 + t in y, one
// double plane at a time through ld (barrier-separated; ends with a barrier).
template <int C>
__device__ __forceinline__ void split_transpose(double *ld, const cd (&x)[16], cd (&y)[16], int r0, int c,
                                                bool to_blocks) {
#pragma unroll
  for (int plane = 0; plane < 2; plane++) {
#pragma unroll
    for (int t = 0; t < 16; t++) {
      const int row = to_blocks ? r0 + 16 * t : 16 * r0 + t;
      ld[slot<C>(row, c)] = plane ? x[t].y : x[t].x;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 16; t++) {
      const int row = to_blocks ? 16 * r0 + t : r0 + 16 * t;
      if (plane)
        y[t].y = ld[slot<C>(row, c)];
      else
        y[t].x = ld[slot<C>(row, c)];
    }
    __syncthreads();
  }
}

// Buffer resource over `words` 32-bit words at p: loads at byte offsets at or
// beyond 4 * words return 0 (the resource's range check). The record count is
// a 32-bit byte count; larger arrays are only ever read below 2^32 bytes.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t word_rsrc(const uint32_t *p, unsigned long long words) {
  const unsigned bytes = words >= 0x3fffffffull ? 0xfffffffcu : (unsigned)(words * 4ull);
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(p), (short)0, (int)bytes, 0x00020000);
}

// Digits of the packing pass, read straight from the operand words. Lane
// (r0, c) takes column m = m0 + c of rows r0 + h L/16 (h < 16) in both halves:
// digit j = row M + m (+ N) starts at bit j b, that is at bit (m b) & 31 of
// word row (M b / 32) + (m b >> 5) (+ N b / 32), as M b and N b are multiples
// of 32. Words at or beyond al read as 0 through the range check. When b does
// not divide 32 a digit may straddle two words; the second comes from a
// resource based one word higher, so every load is one whole dword either in
// or out of range.
struct PackDigits {
  __amdgpu_buffer_rsrc_t w0, w1, u0, u1;
  unsigned off;   // byte offset of the lane's word in row r0
  unsigned step;  // bytes between rows r0 + h L/16 and r0 + (h + 1) L/16
  unsigned sh, bits;
  bool upper;     // the upper half (digits N ..) holds words
  bool one;       // b divides 32: a digit never straddles two words

  template <int L>
  __device__ __forceinline__ void init(const PassArgs &a, int r0, unsigned long long m) {
    bits = (unsigned)a.bits;
    const unsigned long long row_w = (a.M * bits) >> 5, half_w = (a.N * bits) >> 5, mb = m * bits;
    const unsigned long long al = a.al;
    w0 = word_rsrc(a.wa, al);
    w1 = word_rsrc(a.wa + 1, al ? al - 1 : 0);
    // an empty upper half gets empty resources: its loads return 0 without a
    // memory access, and both halves' loads issue together
    upper = half_w < al;
    u0 = word_rsrc(upper ? a.wa + half_w : a.wa, upper ? al - half_w : 0);
    u1 = word_rsrc(upper ? a.wa + half_w + 1 : a.wa, upper ? al - half_w - 1 : 0);
    off = (unsigned)(((unsigned long long)r0 * row_w + (mb >> 5)) * 4ull);
    step = (unsigned)(row_w * (unsigned long long)(L / 16) * 4ull);
    sh = (unsigned)(mb & 31ull);
    one = (bits & (bits - 1u)) == 0u;  // b <= 32 divides 32
  }

  // Lower-half digits only (launches whose operands all leave the upper half
  // empty).
  __device__ __forceinline__ void load_lower(unsigned (&d)[16]) const {
    unsigned lo[16];
#pragma unroll
    for (int h = 0; h < 16; h++) lo[h] = __builtin_amdgcn_raw_buffer_load_b32(w0, off + (unsigned)h * step, 0, 0);
    if (one) {
#pragma unroll
      for (int h = 0; h < 16; h++) d[h] = __builtin_amdgcn_ubfe(lo[h], sh, bits);
    } else {
      unsigned hi[16];
#pragma unroll
      for (int h = 0; h < 16; h++) hi[h] = __builtin_amdgcn_raw_buffer_load_b32(w1, off + (unsigned)h * step, 0, 0);
#pragma unroll
      for (int h = 0; h < 16; h++) d[h] = __builtin_amdgcn_ubfe(__builtin_amdgcn_alignbit(hi[h], lo[h], sh), 0, bits);
    }
  }

  // Lower-half digits into d, upper-half digits into e (0 when empty); every
  // load is issued before the first is used.
  __device__ __forceinline__ void load(unsigned (&d)[16], unsigned (&e)[16]) const {
    unsigned lo[16], ulo[16];
#pragma unroll
    for (int h = 0; h < 16; h++) {
      lo[h] = __builtin_amdgcn_raw_buffer_load_b32(w0, off + (unsigned)h * step, 0, 0);
      ulo[h] = __builtin_amdgcn_raw_buffer_load_b32(u0, off + (unsigned)h * step, 0, 0);
    }
    if (one) {
#pragma unroll
      for (int h = 0; h < 16; h++) {
        d[h] = __builtin_amdgcn_ubfe(lo[h], sh, bits);
        e[h] = __builtin_amdgcn_ubfe(ulo[h], sh, bits);
      }
    } else {
      unsigned hi[16], uhi[16];
#pragma unroll
      for (int h = 0; h < 16; h++) {
        hi[h] = __builtin_amdgcn_raw_buffer_load_b32(w1, off + (unsigned)h * step, 0, 0);
        uhi[h] = __builtin_amdgcn_raw_buffer_load_b32(u1, off + (unsigned)h * step, 0, 0);
      }
#pragma unroll
      for (int h = 0; h < 16; h++) {
        d[h] = __builtin_amdgcn_ubfe(__builtin_amdgcn_alignbit(hi[h], lo[h], sh), 0, bits);
        e[h] = __builtin_amdgcn_ubfe(__builtin_amdgcn_alignbit(uhi[h], ulo[h], sh), 0, bits);
      }
    }
  }
};

// Streaming loads and stores of the sixteen rows K r0 + S t (t < 16) of column
// m0 + c, rows M points apart, for lane (r0, c) of a C-column tile: one lane
// base address, plus the row step S t M, which is the same across the wave
// (scalar registers), so each access costs one address addition and the
// sixteen accesses hold no per-row address registers.
template <int K, int S, bool PLAIN = false>
__device__ __forceinline__ void load_rows16(const cd *base, cd (&v)[16], int r0, int c, unsigned long long m0,
                                            unsigned long long M) {
  const cd *lb = base + (unsigned long long)(K * r0) * M + m0 + c;
#pragma unroll
  for (int t = 0; t < 16; t++) v[t] = ld_pol<PLAIN>(lb + (unsigned long long)(S * t) * M);
}

template <int K, int S, bool PLAIN = false>
__device__ __forceinline__ void store_rows16(cd *base, const cd (&v)[16], int r0, int c, unsigned long long m0,
                                             unsigned long long M) {
  cd *lb = base + (unsigned long long)(K * r0) * M + m0 + c;
#pragma unroll
  for (int t = 0; t < 16; t++) st_pol<PLAIN>(lb + (unsigned long long)(S * t) * M, v[t]);
}

// LOWER (packing pass): every operand of the launch leaves the upper digit
// half empty, so a_{j+N} = 0 and the weighted point is a_j w exactly.
template <int L, int C, bool PACK, int CACHE = 0, bool LOWER = false>
__global__ __launch_bounds__((Tile<L, C>::T)) void k_fwd(PassArgs a) {
  using TC = Tile<L, C>;
  using ST = SplitTile<L, C>;
  bind_product(a);
  // the split transpose pays off where compute dominates (the packing pass);
  // the strided pass keeps the full tile (measured faster at fewer waves)
  constexpr bool SPLIT = ST::ON && PACK;
  constexpr int TILE_BYTES = SPLIT ? ST::BYTES : Tile<L, C>::LDS_ELEMS * 16;
  __shared__ __align__(16) unsigned char lds_raw[TILE_BYTES];
  cd *lds = reinterpret_cast<cd *>(lds_raw);
  const int tid = threadIdx.x;
  const int tile = tile_of_block(a.ntiles);
  const unsigned long long tpb = a.M / C;
  const unsigned long long ob = (unsigned)tile / (unsigned)tpb;  // tpb < 2^31
  const unsigned long long m0 = ((unsigned)tile % (unsigned)tpb) * (unsigned long long)C;
  cd *base = a.data + ob * a.S;
  constexpr int RB0 = TC::FIRSTB;
  constexpr int R0 = 1 << RB0;
  constexpr int PER0 = TC::EPT / R0;
  constexpr int SUB0 = L / R0;
  static_assert(!PACK || !TC::THREE, "packing pass is a power of two");
  // Thread tid owns column c = tid % C and rows r0 + h * L/16, r0 = tid / C,
  // h = 0..15 (T = L*C/16): row r0 + h*L/16 has frequency 16 brev(r0) + brev4(h).
  const int c = tid % C;
  const int r0 = tid / C;
  const unsigned long long m = m0 + c;
  // PACK: digit pairs (a_j, a_{j+N}) of rows r0 + h L/16 and the right-angle
  // weight Omega^{row M + m} = Omega^{r0 M + m} e^{2 pi i h / 64}
  unsigned dre[PACK ? 16 : 1], dim_[PACK ? 16 : 1];
  cd wpack = make_double2(1.0, 0.0);
  cd tb0[4];  // split tile: stage-0 level twiddles, loaded ahead of the digits
  if constexpr (SPLIT) level_bases<4>(a.tl, r0, 1, tb0);
  if constexpr (PACK) {
    PackDigits pd;
    pd.template init<L>(a, r0, m);
    if constexpr (LOWER)
      pd.load_lower(dre);
    else
      pd.load(dre, dim_);
    wpack = omega_direct(a.roots, (unsigned long long)r0 * a.M + m);
  }
  if constexpr (SPLIT) {
    // PER0 = 1, SUB0 = 16: rows r0 + 16 t, then rows 16 r0 + t after the transpose
    cd x[16], y[16];
#pragma unroll
    for (int t = 0; t < 16; t++) {
      if constexpr (PACK && LOWER) {
        const cd w = cmul(wpack, w64(t));
        const double r = (double)dre[t];
        x[t] = make_double2(r * w.x, r * w.y);
      } else if constexpr (PACK) {
        // (a_j + i a_{j+N}) w (a_{j+N} = 0 throughout when the upper half is empty)
        x[t] = cmul(make_double2((double)dre[t], (double)dim_[t]), cmul(wpack, w64(t)));
      } else {
        x[t] = ld_nt(&base[(unsigned long long)(r0 + 16 * t) * a.M + m]);
      }
    }
    dif_levels<4>(x, tb0);
    split_transpose<C>(reinterpret_cast<double *>(lds_raw), x, y, r0, c, false);
    dif_levels_q0<4>(y);
    // row 16 r0 + t has frequency 16 brev4(t) + brev4(r0)
    Walk16 w;
    if (a.M > 1) w.template init_shared<C>(a.roots, a.tw_mul * m * brev((unsigned)r0, 4), a.tw_mul * m * 16ull);
    if (a.M > 1) {
#pragma unroll
      for (int t = 0; t < 16; t++) y[t] = cmul(y[t], w.at(t));
    }
    store_rows16<16, 1>(base, y, r0, c, m0, a.M);
    return;
  }
#pragma unroll
  for (int i = 0; i < PER0; i++) {
    int q = r0 + i * (TC::T / C);
    cd x[R0];
#pragma unroll
    for (int t = 0; t < R0; t++) {
      unsigned long long row = (unsigned long long)(q + t * SUB0);
      if constexpr (PACK && LOWER) {
        const int h = i + t * PER0;
        const cd w = cmul(wpack, w64(h));
        x[t] = make_double2((double)dre[h] * w.x, (double)dre[h] * w.y);
      } else if constexpr (PACK) {
        const int h = i + t * PER0;
        x[t] = cmul(make_double2((double)dre[h], (double)dim_[h]), cmul(wpack, w64(h)));
      } else {
        x[t] = ld_pol<(CACHE & 1) != 0>(&base[row * a.M + m]);
      }
    }
    dif_levels<RB0>(x, a.tl, q, 1);
#pragma unroll
    for (int t = 0; t < R0; t++) lds[slot<C>(q + t * SUB0, c)] = x[t];
  }
  __syncthreads();
  forward_rest<L, C>(lds, a.tl, tid);
  if constexpr (TC::THREE) {
    for (int e = tid; e < TC::E; e += TC::T) {
      int row = e / C, cc = e % C;
      unsigned long long mm = m0 + cc;
      cd v = lds[pad16(e)];
      if (a.M > 1) {
        unsigned long long k = freq_of_row<L>((unsigned)row);
        v = cmul(v, omega_pow(a.roots, mm * k * a.tw_mul));
      }
      base[(unsigned long long)row * a.M + mm] = v;
    }
  } else {
    // pass twiddle e^{2 pi i m k / S} = Omega^{tw_mul m k}
    Walk16 w;
    if (a.M > 1) w.template init_shared<C>(a.roots, a.tw_mul * m * 16ull * brev((unsigned)r0, TC::LOG - 4), a.tw_mul * m);
#pragma unroll
    for (int h = 0; h < 16; h++) {
      const int row = r0 + h * (L / 16);
      cd v = lds[slot<C>(row, c)];
      if (a.M > 1) v = cmul(v, w.at(h));
      st_pol<(CACHE & 2) != 0>(&base[(unsigned long long)row * a.M + m], v);
    }
  }
}

// ------------------------------------------------------------ inverse pass

// Mirror of k_fwd: staged load with the conjugate pass twiddle, DIT stages in
// LDS, stage 0 stores straight to global, or rounds to coefficients: MODE 0
// stores points, MODE 1 writes the 2N coefficients, MODE 2 their group sums
// (a.ybuf). The two unpacking modes are separate kernels, so the group mode
// is not held to the registers of the coefficient stores.
// The unpacking pass of a split tile is held to 168 VGPRs (three waves per
// SIMD; measured faster than two waves without the two spilled registers).
template <int L, int C, int MODE, int CACHE = 0>
__global__ __launch_bounds__((Tile<L, C>::T))
__attribute__((amdgpu_waves_per_eu((MODE == 1 && SplitTile<L, C>::ON && C >= 8) ? 3 : 1, 8))) void k_inv(PassArgs a) {
  constexpr bool UNPACK = MODE != 0;
  constexpr bool GROUPS = MODE == 2;
  using TC = Tile<L, C>;
  using ST = SplitTile<L, C>;
  bind_product(a);
  __shared__ __align__(16) unsigned char lds_raw[ST::BYTES];
  cd *lds = reinterpret_cast<cd *>(lds_raw);
  const int tid = threadIdx.x;
  const int tile = tile_of_block(a.ntiles);
  const unsigned long long tpb = a.M / C;
  const unsigned long long ob = (unsigned)tile / (unsigned)tpb;  // tpb < 2^31
  const unsigned long long m0 = ((unsigned)tile % (unsigned)tpb) * (unsigned long long)C;
  cd *base = a.data + ob * a.S;
  static_assert(!UNPACK || !TC::THREE, "unpacking pass is a power of two");
  const int c = tid % C;
  const int r0 = tid / C;
  const unsigned long long m = m0 + c;
  if constexpr (ST::ON) {
    // rows 16 r0 + t (frequency 16 brev4(t) + brev4(r0)) with the conjugate
    // pass twiddle, stage 1 in registers, transpose to rows r0 + 16 t, stage 0
    cd x[16], y[16], tb0[4];
    load_rows16<16, 1, (CACHE & 1) != 0>(base, y, r0, c, m0, a.M);
    Walk16 w;
    if (a.M > 1) {
      w.template init_shared<C>(a.roots, a.tw_mul * m * brev((unsigned)r0, 4), a.tw_mul * m * 16ull);
#pragma unroll
      for (int t = 0; t < 16; t++) y[t] = cmulc(y[t], w.at(t));
    }
    dit_levels_q0<4>(y);
    // stage-0 level twiddles, loaded while the transpose runs
    level_bases<4>(a.tl, r0, 1, tb0);
    double *ld = reinterpret_cast<double *>(lds_raw);
    split_transpose<C>(ld, y, x, r0, c, true);
    dit_levels<4>(x, tb0);
    if constexpr (!UNPACK) {
      store_rows16<1, 16, (CACHE & 2) != 0>(base, x, r0, c, m0, a.M);
    } else {
      double worst = 0.0;
      const cd wun = omega_direct(a.roots, (unsigned long long)r0 * a.M + m);
      if constexpr (GROUPS) {
        // group mode (coefficients below 2^51): the real coefficients, then
        // the imaginary ones, each summed per row from LDS
        long long *lc = reinterpret_cast<long long *>(lds_raw);
        long long ci[16];
#pragma unroll
        for (int t = 0; t < 16; t++) {
          const cd f = cmulc(x[t], cmul(wun, w64(t)));
          lc[slot<C>(r0 + 16 * t, c)] = round_coef(f.x * a.inv_n, true, worst);
          ci[t] = round_coef(f.y * a.inv_n, true, worst);
        }
        __syncthreads();
        write_group_half<L, C>(a, lc, m0, tid, 0);
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 16; t++) lc[slot<C>(r0 + 16 * t, c)] = ci[t];
        __syncthreads();
        write_group_half<L, C>(a, lc, m0, tid, 1);
      } else {
#pragma unroll
        for (int t = 0; t < 16; t++)
          worst = fmax(worst, round_store(a, cmulc(x[t], cmul(wun, w64(t))), (unsigned long long)(r0 + 16 * t), m, L));
      }
      margin_reduce(a.margin, worst, a.margin_mask);
    }
    return;
  }
  if constexpr (TC::THREE) {
    for (int e = tid; e < TC::E; e += TC::T) {
      int row = e / C, cc = e % C;
      unsigned long long mm = m0 + cc;
      cd v = base[(unsigned long long)row * a.M + mm];
      if (a.M > 1) {
        unsigned long long k = freq_of_row<L>((unsigned)row);
        v = cmulc(v, omega_pow(a.roots, mm * k * a.tw_mul));
      }
      lds[pad16(e)] = v;
    }
  } else {
    // all sixteen loads in flight first, then the conjugate pass twiddle
    // (rows r0 + h * L/16 as in k_fwd) on the way into LDS
    cd xs[16];
    load_rows16<1, L / 16, (CACHE & 1) != 0>(base, xs, r0, c, m0, a.M);
    Walk16 w;
    if (a.M > 1) w.template init_shared<C>(a.roots, a.tw_mul * m * 16ull * brev((unsigned)r0, TC::LOG - 4), a.tw_mul * m);
#pragma unroll
    for (int h = 0; h < 16; h++) {
      const int row = r0 + h * (L / 16);
      lds[slot<C>(row, c)] = (a.M > 1) ? cmulc(xs[h], w.at(h)) : xs[h];
    }
  }
  __syncthreads();
  inverse_front<L, C>(lds, a.tl, tid);
  constexpr int RB0 = TC::FIRSTB;
  constexpr int R0 = 1 << RB0;
  constexpr int PER0 = TC::EPT / R0;
  constexpr int SUB0 = L / R0;
  double worst = 0.0;
  cd wun = make_double2(1.0, 0.0);
  if (UNPACK) wun = omega_direct(a.roots, (unsigned long long)r0 * a.M + m);
  constexpr bool groups = GROUPS;
#pragma unroll
  for (int i = 0; i < PER0; i++) {
    int q = r0 + i * (TC::T / C);
    cd x[R0];
#pragma unroll
    for (int t = 0; t < R0; t++) x[t] = lds[slot<C>(q + t * SUB0, c)];
    dit_levels<RB0>(x, a.tl, q, 1);
#pragma unroll
    for (int t = 0; t < R0; t++) {
      unsigned long long row = (unsigned long long)(q + t * SUB0);
      if constexpr (UNPACK) {
        const int h = i + t * PER0;
        const cd f = cmulc(x[t], cmul(wun, w64(h)));
        if (groups) {
          // in place: this thread alone reads and writes these slots in the final stage
          // group mode implies coefficients below 2^51
          const long long cr = round_coef(f.x * a.inv_n, true, worst);
          const long long ci = round_coef(f.y * a.inv_n, true, worst);
          reinterpret_cast<longlong2 *>(lds)[slot<C>(q + t * SUB0, c)] = make_longlong2(cr, ci);
        } else {
          worst = fmax(worst, round_store(a, f, row, m, L));
        }
      } else {
        st_pol<(CACHE & 2) != 0>(&base[row * a.M + m], x[t]);
      }
    }
  }
  if constexpr (UNPACK) {
    if constexpr (groups) {
      __syncthreads();
      write_group_sums<L, C>(a, lds, m0, tid);
    }
    margin_reduce(a.margin, worst, a.margin_mask);
  }
}

// ------------------------------------------------------------- middle pass

// Forward DIF of one contiguous block of L points, ending with the innermost
// stage's outputs in registers: xr[i * R + t] for butterfly i of that stage.
template <int L, bool PACK, int CACHE = 0>
__device__ __forceinline__ void mid_forward(const PassArgs &a, cd *lds, const cd *__restrict__ src,
                                            const uint32_t *__restrict__ w, unsigned long long wl,
                                            unsigned long long blk0, cd (&xr)[Tile<L, 1>::EPT], int tid) {
  using TC = Tile<L, 1>;
  constexpr int RB0 = TC::FIRSTB;
  constexpr int R0 = 1 << RB0;
  constexpr int PER0 = TC::EPT / R0;
  constexpr int SUB0 = L / R0;
  // Stage 0 straight from global.
#pragma unroll
  for (int i = 0; i < PER0; i++) {
    int q = tid + i * TC::T;
    cd x[R0];
#pragma unroll
    for (int t = 0; t < R0; t++) {
      unsigned long long j = blk0 + (unsigned long long)(q + t * SUB0);
      x[t] = PACK ? packed_point(a, w, wl, j) : ld_pol<(CACHE & 1) != 0>(&src[j]);
    }
    dif_levels<RB0>(x, a.tl, q, 1);
    if (TC::NST2 == 1 && !TC::THREE) {
#pragma unroll
      for (int t = 0; t < R0; t++) xr[i * R0 + t] = x[t];
    } else {
#pragma unroll
      for (int t = 0; t < R0; t++) lds[slot_step<SUB0>(q, t)] = x[t];
    }
  }
  if (TC::NST2 == 1 && !TC::THREE) return;
  __syncthreads();
  // Middle stages in LDS.
  int span = L >> RB0;
  const int last = TC::THREE ? TC::NST2 : TC::NST2 - 1;  // stages run in LDS
  for (int s = 1; s < last; s++) {
    int rb = stage_bits<L, 1>(s);
    lds_stage_any<L, 1, false>(lds, rb, span, a.tl, tid);
    __syncthreads();
    span >>= rb;
  }
  if (TC::THREE) {
    constexpr int PER = TC::EPT / 3;
#pragma unroll
    for (int i = 0; i < PER; i++) {
      int b3 = tid + i * TC::T;
      cd p = lds[pad16(3 * b3)], q = lds[pad16(3 * b3 + 1)], r = lds[pad16(3 * b3 + 2)];
      dft3<false>(p, q, r);
      xr[3 * i] = p;
      xr[3 * i + 1] = q;
      xr[3 * i + 2] = r;
    }
  } else {
    constexpr int RBL = TC::MAXB;  // last stage is a full MAXB stage when NST2 > 1
    constexpr int R = 1 << RBL;
    constexpr int PER = TC::EPT / R;
    // span == R here: each butterfly is one contiguous block of R rows
#pragma unroll
    for (int i = 0; i < PER; i++) {
      int blk = tid + i * TC::T;
      cd x[R];
#pragma unroll
      for (int t = 0; t < R; t++) x[t] = lds[slot<R>(blk, t)];
      dif_levels_q0<RBL>(x);
#pragma unroll
      for (int t = 0; t < R; t++) xr[i * R + t] = x[t];
    }
  }
}

// Last forward pass of A (and B), pointwise product (or square), first inverse
// pass on one contiguous block of L points. Ends with the outermost inverse
// stage's outputs in registers: out[i * R0 + t] holds row q + t * (L / R0),
// q = tid + i * T (natural order). The caller stores or unpacks them.
template <int L, bool SQUARE, bool PACK, int CACHE = 0>
__device__ __forceinline__ void mid_pipeline(const PassArgs &a, cd *lds, unsigned long long blk0,
                                             cd (&xa)[Tile<L, 1>::EPT], int tid) {
  using TC = Tile<L, 1>;
  mid_forward<L, PACK, CACHE>(a, lds, a.data, a.wa, a.al, blk0, xa, tid);
  if (!SQUARE) {
    cd xb[TC::EPT];
    __syncthreads();
    mid_forward<L, PACK, CACHE>(a, lds, a.data_b, a.wb, a.bl, blk0, xb, tid);
#pragma unroll
    for (int i = 0; i < TC::EPT; i++) xa[i] = cmul(xa[i], xb[i]);
  } else {
#pragma unroll
    for (int i = 0; i < TC::EPT; i++) xa[i] = cmul(xa[i], xa[i]);
  }
  constexpr int RB0 = TC::FIRSTB;
  constexpr int R0 = 1 << RB0;
  constexpr int PER0 = TC::EPT / R0;
  constexpr int SUB0 = L / R0;
  if (!(TC::NST2 == 1 && !TC::THREE)) {
    __syncthreads();  // every thread is done reading B's innermost stage from LDS
    if (TC::THREE) {
      constexpr int PER = TC::EPT / 3;
#pragma unroll
      for (int i = 0; i < PER; i++) {
        int b3 = tid + i * TC::T;
        cd p = xa[3 * i], q = xa[3 * i + 1], r = xa[3 * i + 2];
        dft3<true>(p, q, r);
        lds[pad16(3 * b3)] = p;
        lds[pad16(3 * b3 + 1)] = q;
        lds[pad16(3 * b3 + 2)] = r;
      }
    } else {
      constexpr int RBL = TC::MAXB;
      constexpr int R = 1 << RBL;
      constexpr int PER = TC::EPT / R;
#pragma unroll
      for (int i = 0; i < PER; i++) {
        int blk = tid + i * TC::T;
        cd x[R];
#pragma unroll
        for (int t = 0; t < R; t++) x[t] = xa[i * R + t];
        dit_levels_q0<RBL>(x);
#pragma unroll
        for (int t = 0; t < R; t++) lds[slot<R>(blk, t)] = x[t];
      }
    }
    __syncthreads();
    // remaining inverse stages in LDS, innermost to outermost, excluding stage 0
    int first_lds = TC::THREE ? TC::NST2 - 1 : TC::NST2 - 2;
    int span = TC::THREE ? 3 : 1;
    for (int s = TC::NST2 - 1; s > first_lds; s--) span <<= stage_bits<L, 1>(s);
    for (int s = first_lds; s >= 1; s--) {
      int rb = stage_bits<L, 1>(s);
      span <<= rb;
      lds_stage_any<L, 1, true>(lds, rb, span, a.tl, tid);
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < PER0; i++) {
    int q = tid + i * TC::T;
    cd x[R0];
    if (TC::NST2 == 1 && !TC::THREE) {
#pragma unroll
      for (int t = 0; t < R0; t++) x[t] = xa[i * R0 + t];
    } else {
#pragma unroll
      for (int t = 0; t < R0; t++) x[t] = lds[slot_step<SUB0>(q, t)];
    }
    dit_levels<RB0>(x, a.tl, q, 1);
#pragma unroll
    for (int t = 0; t < R0; t++) xa[i * R0 + t] = x[t];
  }
}

// Middle pass of a multi-pass plan (in place into a.data), or a whole one-pass
// product rounded into a.coef when PACK/UNPACK. BPW contiguous blocks per
// workgroup (each with its own LDS tile): blocks of L < 1024 (1536 for 3 2^k)
// have fewer than 64 threads, and one block per workgroup would leave most
// lanes of the wave idle.
template <int L, bool SQUARE, bool PACK, bool UNPACK, int BPW = 1, int CACHE = 0>
__global__ __launch_bounds__((Tile<L, 1>::T * BPW)) void k_mid(PassArgs a) {
  using TC = Tile<L, 1>;
  static_assert(BPW == 1 || !PACK, "one-pass products take one workgroup each");
  bind_product(a);
  __shared__ cd lds_all[TC::LDS_ELEMS * BPW];
  const int sub = BPW > 1 ? (int)threadIdx.x / TC::T : 0;
  const int tid = BPW > 1 ? (int)threadIdx.x % TC::T : (int)threadIdx.x;
  cd *lds = lds_all + sub * TC::LDS_ELEMS;
  const unsigned long long blk0 = ((unsigned long long)blockIdx.x * BPW + sub) * L;
  cd xa[TC::EPT];
  mid_pipeline<L, SQUARE, PACK, CACHE>(a, lds, blk0, xa, tid);
  constexpr int R0 = 1 << TC::FIRSTB;
  constexpr int PER0 = TC::EPT / R0;
  constexpr int SUB0 = L / R0;
  double worst = 0.0;
#pragma unroll
  for (int i = 0; i < PER0; i++) {
    int q = tid + i * TC::T;
#pragma unroll
    for (int t = 0; t < R0; t++) {
      unsigned long long j = blk0 + (unsigned long long)(q + t * SUB0);
      if (UNPACK)
        worst = fmax(worst, unpack_point(a, xa[i * R0 + t], j));
      else
        st_pol<(CACHE & 2) != 0>(&a.data[j], xa[i * R0 + t]);
    }
  }
  if (UNPACK) margin_reduce(a.margin, worst, a.margin_mask);
}

// ------------------------------------------------------------ carry resolve

// Carry functions of single-bit carry chains: bit0 = carry out for carry in 0,
// bit1 = carry out for carry in 1. compose(f, g) applies f first, then g.
__device__ __forceinline__ unsigned carry_compose(unsigned f, unsigned g) {
  unsigned h0 = (f & 1u) ? (g >> 1) & 1u : g & 1u;
  unsigned h1 = (f & 2u) ? (g >> 1) & 1u : g & 1u;
  return h0 | (h1 << 1);
}


// Product words from rounded coefficients: value = sum_j c_j 2^{b j}.
// Word w collects P_w = sum_j piece(c_j 2^{bj}, word w) (< 2^37); then
// v_w = lo32(P_w) + hi(P_{w-1}) and a carry chain of single bits resolves v.
// Each thread owns CWPT consecutive words; a block owns CWPT * CT words.
constexpr int CWPT = 4;
constexpr int CT = 256;

struct CarryArgs {
  const MulDesc *desc;       // batched: product blockIdx.y (local) / blockIdx.x (fix), else nullptr
  unsigned long long coef_stride;  // batched: coefficients per product (2N)
  const long long *coef;
  unsigned long long ncoef;
  int bits;
  int cbits;                 // every coefficient < 2^cbits (cbits <= 63)
  uint32_t *out;
  unsigned long long CL;
  unsigned char *blk_fn;     // per block: bit0 = carry out if carry in 0, bit1 = if carry in 1
  unsigned long long nblk;
  const uint32_t *add;       // optional addend words (value += add), addl words
  unsigned long long addl;
  int coef_T;                // coefficient j at coef[(j % cM) * crows + j / cM]
  unsigned cM, crows;
  unsigned lds_slots;        // int64 slots of dynamic LDS per block (coefficient window)
  const uint32_t *ybuf;      // group mode: group sums instead of coefficients (ncoef digits in all)
  int gC, gK, gKs;           // digits per group, words per group sum, group stride (words)
  long long ng;              // groups: ncoef / gC
  unsigned wpb;              // words per carry block
  unsigned long long ystride;  // batched group mode: group-sum words per product
};

__device__ __forceinline__ long long coef_at(const CarryArgs &a, long long j) {
  if (!a.coef_T) return a.coef[j];
  unsigned jj = (unsigned)j;
  return a.coef[(unsigned long long)(jj % a.cM) * a.crows + jj / a.cM];
}

__device__ __forceinline__ void word_partials(const CarryArgs &a, long long wfirst, unsigned long long (&P)[CWPT + 1]) {
  // words wfirst-1 .. wfirst+CWPT-1 -> P[0..CWPT]
#pragma unroll
  for (int k = 0; k <= CWPT; k++) P[k] = 0;
  long long lo_bit = 32 * (wfirst - 1) - a.cbits + 1;
  long long jlo = lo_bit <= 0 ? 0 : (lo_bit + a.bits - 1) / a.bits;
  long long hi_bit = 32 * (wfirst + CWPT) - 1;
  long long jhi = hi_bit / a.bits;
  if (jhi >= (long long)a.ncoef) jhi = (long long)a.ncoef - 1;
  for (long long j = jlo; j <= jhi; j++) {
    unsigned long long c = (unsigned long long)coef_at(a, j);
    if (!c) continue;
    long long shift = (long long)a.bits * j - 32 * (wfirst - 1);
#pragma unroll
    for (int k = 0; k <= CWPT; k++) {
      long long sh = shift - 32 * k;
      unsigned long long piece;
      if (sh >= 0)
        piece = sh < 32 ? (unsigned long long)(uint32_t)(c << sh) : 0ull;
      else
        piece = -sh < 64 ? (unsigned long long)(uint32_t)(c >> (-sh)) : 0ull;
      P[k] += piece;
    }
  }
  if (a.add) {
#pragma unroll
    for (int k = 0; k <= CWPT; k++) {
      long long w = wfirst - 1 + k;
      if (w >= 0 && w < (long long)a.addl) P[k] += a.add[w];
    }
  }
}

// Batched carries: product p's coefficients at coef + p * coef_stride, its
// block functions at blk_fn + p * nblk, its words and sizes from desc[p].
__device__ __forceinline__ void bind_carry(CarryArgs &a, unsigned long long p) {
  if (!a.desc) return;
  const MulDesc &d = a.desc[p];
  if (a.coef) a.coef += p * a.coef_stride;
  if (a.ybuf) {
    a.ybuf += p * a.ystride;
    a.gK = d.gk;
  }
  a.bits = d.bits;
  a.cbits = d.cbits;
  a.out = d.c;
  a.CL = d.cl;
  a.add = d.d;
  a.addl = d.dl;
  a.blk_fn += p * a.nblk;
  a.nblk = (d.cl + a.wpb - 1) / a.wpb;
}

// Coefficient window of one carry block in LDS: words [W0, W0 + CT*CWPT) need
// coefficients j with b j in (32 (W0-1) - cbits, 32 (W0 + CT*CWPT)); the
// window is sized per launch (dynamic LDS) from the smallest digit size.
// 0 (no staging: coefficients are read from global memory) when the window
// would exceed CARRY_LDS_MAX slots (digit sizes below 4 bits).
constexpr unsigned CARRY_LDS_MAX = 8256;
__host__ __device__ inline unsigned carry_slots(int bits, int cbits) {
  unsigned n = (unsigned)((32ull * CT * CWPT + (unsigned long long)cbits + 64) / (unsigned long long)bits + 4);
  return n <= CARRY_LDS_MAX ? n : 0u;
}

// floor(x / b) for 0 <= x < 2^40 through a double reciprocal plus one fix-up.
__device__ __forceinline__ long long div_small(long long x, int b, double inv_b) {
  long long q = (long long)((double)x * inv_b);
  if ((q + 1) * b <= x) q++;
  if (q * b > x) q--;
  return q;
}

constexpr int CARRY_GCOV = 4;  // most group sums covering one word

// x / b for 0 <= x < 2^24 and 1 <= b <= 64 through the float reciprocal,
// corrected to the exact floor.
__device__ __forceinline__ int fdiv_small(int x, int b, float inv_b) {
  int q = (int)((float)x * inv_b);
  if (q * b > x) q--;
  else if ((q + 1) * b <= x) q++;
  return q;
}

// P_w of a block word from the staged window win[0..cnt) (coefficient jbase +
// i at win[i]) in block-relative 32-bit arithmetic: x = 32 (w - W0) + D is the
// word's lowest bit above bit b jbase (D = 32 W0 - b jbase, small), so every
// numerator stays far below 2^24 and the float-reciprocal division is exact.
__device__ __forceinline__ unsigned long long partial_rel(const long long *win, int cnt, int bits, int cbits,
                                                          float inv_b, int x) {
  const int lo = x - cbits + 1;
  const int ilo = lo <= 0 ? 0 : fdiv_small(lo + bits - 1, bits, inv_b);
  int ihi = x + 31 < 0 ? -1 : fdiv_small(x + 31, bits, inv_b);
  if (ihi >= cnt) ihi = cnt - 1;
  unsigned long long P = 0;
  int sh = bits * ilo - x;  // in (-cbits, 32)
  for (int i = ilo; i <= ihi; i++, sh += bits) {
    const unsigned long long c = (unsigned long long)win[i];
    P += (unsigned long long)(uint32_t)(sh >= 0 ? c << sh : c >> -sh);
  }
  return P;
}

// Words of one block from coefficients: P_w (< 2^37) per word, v_w = lo(P_w) +
// hi(P_{w-1}), a wave-shuffle scan of per-thread carry functions, words
// written for block carry-in 0 and the block's function saved for k_carry_fix.
// The coefficient window is staged into LDS with eight loads in flight per
// thread; each thread's word w-1 partial comes from its left neighbour.
static __global__ __launch_bounds__(CT) void k_carry_local(CarryArgs a) {
  extern __shared__ long long cl[];
  __shared__ unsigned wave_fn[CT / 64];
  const int tid = threadIdx.x;
  bind_carry(a, blockIdx.y);
  if (blockIdx.x >= a.nblk) return;
  const long long W0 = (long long)blockIdx.x * CT * CWPT;
  const long long wfirst = W0 + (long long)tid * CWPT;
  const double inv_b = 1.0 / (double)a.bits;
  long long lo_bit = 32 * (W0 - 1) - a.cbits + 1;
  long long jbase = lo_bit <= 0 ? 0 : div_small(lo_bit + a.bits - 1, a.bits, inv_b);
  long long jend = div_small(32 * (W0 + (long long)CT * CWPT) - 1, a.bits, inv_b);  // inclusive
  if (jend >= (long long)a.ncoef) jend = (long long)a.ncoef - 1;
  const long long cnt = jend >= jbase ? jend - jbase + 1 : 0;
  const bool staged = cnt <= (long long)a.lds_slots;
  if (staged) {
    for (long long t0 = 0; t0 < cnt; t0 += 8 * CT) {
      long long r[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const long long t = t0 + u * CT + tid;
        r[u] = t < cnt ? coef_at(a, jbase + t) : 0;
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const long long t = t0 + u * CT + tid;
        if (t < cnt) cl[t] = r[u];
      }
    }
  }
  __syncthreads();
  unsigned long long P[CWPT + 1];
  if (staged) {
    const int lane = tid & 63;
    const int D = (int)(32 * W0 - (long long)a.bits * jbase);
    const float inv_bf = 1.0f / (float)a.bits;
    const int cnti = (int)cnt, x0 = 32 * tid * CWPT + D;  // lowest bit of word wfirst
#pragma unroll
    for (int k = 1; k <= CWPT; k++) {
      long long w = wfirst - 1 + k;
      P[k] = (w < (long long)a.CL) ? partial_rel(cl, cnti, a.bits, a.cbits, inv_bf, x0 + 32 * (k - 1)) : 0ull;
    }
    // word wfirst-1 is the left neighbour's last word
    unsigned long long left = __shfl_up(P[CWPT], 1, 64);
    if (lane == 0)
      left = (wfirst >= 1 && wfirst - 1 < (long long)a.CL) ? partial_rel(cl, cnti, a.bits, a.cbits, inv_bf, x0 - 32)
                                                          : 0ull;
    P[0] = left;
    if (a.add) {
#pragma unroll
      for (int k = 0; k <= CWPT; k++) {
        long long w = wfirst - 1 + k;
        if (w >= 0 && w < (long long)a.addl && w < (long long)a.CL) P[k] += a.add[w];
      }
    }
  } else {
    if (wfirst < (long long)a.CL)
      word_partials(a, wfirst, P);
    else
      for (int k = 0; k <= CWPT; k++) P[k] = 0;
  }
  if (wfirst == 0) P[0] = 0;
  unsigned long long v[CWPT];
#pragma unroll
  for (int k = 0; k < CWPT; k++) v[k] = (P[k + 1] & 0xffffffffull) + (P[k] >> 32);
  unsigned f0 = 0, f1 = 1;
#pragma unroll
  for (int k = 0; k < CWPT; k++) {
    f0 = (unsigned)((v[k] + f0) >> 32);
    f1 = (unsigned)((v[k] + f1) >> 32);
  }
  const unsigned f = f0 | (f1 << 1);
  // inclusive scan of carry functions within the wave, then across waves
  const int lane = tid & 63, wave = tid >> 6;
  unsigned inc = f;
  for (int off = 1; off < 64; off <<= 1) {
    unsigned gq = __shfl_up(inc, off, 64);
    if (lane >= off) inc = carry_compose(gq, inc);
  }
  unsigned excl = __shfl_up(inc, 1, 64);
  if (lane == 0) excl = 2u;
  if (lane == 63) wave_fn[wave] = inc;
  __syncthreads();
  unsigned pre = 2u;
  for (int w = 0; w < wave; w++) pre = carry_compose(pre, wave_fn[w]);
  if (tid == 0) {
    unsigned tot = 2u;
    for (int w = 0; w < CT / 64; w++) tot = carry_compose(tot, wave_fn[w]);
    a.blk_fn[blockIdx.x] = (unsigned char)tot;
  }
  unsigned cin = carry_compose(pre, excl) & 1u;
  uint32_t o[CWPT];
#pragma unroll
  for (int k = 0; k < CWPT; k++) {
    unsigned long long sum = v[k] + cin;
    o[k] = (uint32_t)sum;
    cin = (unsigned)(sum >> 32);
  }
  if (wfirst + CWPT <= (long long)a.CL && ((reinterpret_cast<uintptr_t>(a.out + wfirst) & 15) == 0)) {
    *reinterpret_cast<uint4 *>(a.out + wfirst) = make_uint4(o[0], o[1], o[2], o[3]);
  } else {
#pragma unroll
    for (int k = 0; k < CWPT; k++)
      if (wfirst + k < (long long)a.CL) a.out[wfirst + k] = o[k];
  }
}

// Carry resolve from group sums (see write_group_sums): group g (digits
// [G g, G g + G)) covers words W_g .. W_g + K - 1, W_g = floor(G b g / 32),
// and at most CARRY_GCOV groups cover one word. A block owns GCT * GWPT
// words; the group sums under them are staged in LDS with 16-byte loads;
// each thread walks its GWPT + 1 words (the first is its left neighbour's
// last) with a running group index, forms P_w, v_w = lo(P_w) + hi(P_{w-1})
// and the carry-function scan of k_carry_local.
constexpr int GCT = 256;
constexpr int GWPT = 8;  // 2048-word blocks: 17 KB of staged groups, nine blocks per CU
template <int GCOV>  // group sums that can cover one word (<= CARRY_GCOV)
static __global__ __launch_bounds__(GCT) void k_carry_groups(CarryArgs a) {
  extern __shared__ uint32_t gl_w[];
  __shared__ unsigned wave_fn[GCT / 64];
  bind_carry(a, blockIdx.y);  // batched: product blockIdx.y
  if ((unsigned long long)blockIdx.x >= a.nblk) return;
  const int tid = threadIdx.x;
  const long long W0 = (long long)blockIdx.x * GCT * GWPT;
  const int gb = a.gC * a.bits;  // bits per group, >= 32
  const double inv_gb = 1.0 / (double)gb;
  const long long ng = a.ng;
  long long g_first = W0 == 0 ? 0 : div_small(32 * (W0 - 1) + 31, gb, inv_gb) - (CARRY_GCOV - 1);
  if (g_first < 0) g_first = 0;
  long long g_last = div_small(32 * (W0 + (long long)GCT * GWPT - 1) + 31, gb, inv_gb);
  if (g_last >= ng) g_last = ng - 1;
  // LDS word i lives at i + i / 32: threads read words GWPT apart, and the pad
  // keeps those reads on distinct banks
  {
    const int nvec = g_last >= g_first ? (int)((g_last - g_first + 1) * a.gKs / 4) : 0;
    const uint4 *src = reinterpret_cast<const uint4 *>(a.ybuf + g_first * a.gKs);
    for (int t0 = 0; t0 < nvec; t0 += 8 * GCT) {
      uint4 r[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int t = t0 + u * GCT + tid;
        r[u] = t < nvec ? src[t] : make_uint4(0u, 0u, 0u, 0u);
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int t = t0 + u * GCT + tid;
        if (t < nvec) {
          const unsigned i = (unsigned)(4 * t), pi = i + (i >> 5);  // 4 words never straddle a pad
          gl_w[pi] = r[u].x;
          gl_w[pi + 1] = r[u].y;
          gl_w[pi + 2] = r[u].z;
          gl_w[pi + 3] = r[u].w;
        }
      }
    }
  }
  __syncthreads();
  const long long wfirst = W0 + (long long)tid * GWPT;
  // Block-relative 32-bit arithmetic: staged group gl (group g_first + gl)
  // starts at bit B0 + gl gb, B0 = g_first gb, i.e. in block word
  // D + ((r0 + gl gb) >> 5), r0 = B0 & 31, D = (B0 >> 5) - W0 (<= 0).
  const unsigned long long B0 = (unsigned long long)g_first * (unsigned long long)gb;
  const int r0 = (int)(B0 & 31ull);
  const int D = (int)((long long)(B0 >> 5) - W0);
  const int ngl = (int)(g_last - g_first + 1);
  // words below lim (relative) exist: w < CL
  const long long lim64 = (long long)a.CL - W0;
  const int lim = lim64 > (1 << 30) ? (1 << 30) : (int)lim64;
  const int wmin = W0 == 0 ? 0 : -1;  // word W0 - 1 exists unless W0 = 0
  unsigned long long P[GWPT + 1];
  if ((gb & 31) == 0) {
    // Word-aligned group starts (G b a multiple of 32, e.g. b even): group gl
    // starts at block word D + gl Q, so word wl = D + gh Q + o takes word
    // o + t Q of group gh - t, at LDS index ih - t (gKs - Q), ih = gh gKs + o;
    // all advance by one per word. o < Q < K, so term 0 only needs gh >= 0,
    // and gh < ngl for every word of the block.
    const int Q = gb >> 5, dg = a.gKs - Q;
    int wl = tid * GWPT - 1;
    const int x0 = wl - D;  // >= -1
    int gh = x0 < 0 ? -1 : (int)((unsigned)x0 / (unsigned)Q);
    int o = x0 - gh * Q;
    int ih = gh * a.gKs + o;
#pragma unroll
    for (int k = 0; k <= GWPT; k++, wl++) {
      if (k > 0) {
        const bool wrap = o + 1 == Q;
        o = wrap ? 0 : o + 1;
        gh += wrap ? 1 : 0;
        ih += wrap ? 1 + dg : 1;
      }
      const unsigned i0 = gh >= 0 ? (unsigned)ih : 0u;
      unsigned long long p = gh >= 0 ? gl_w[i0 + (i0 >> 5)] : 0u;
#pragma unroll
      for (int t = 1; t < GCOV; t++) {
        const bool ok = gh >= t && o + t * Q < a.gK;
        const unsigned i = ok ? (unsigned)(ih - t * dg) : 0u;
        const unsigned x = gl_w[i + (i >> 5)];
        p += ok ? x : 0u;
      }
      P[k] = (wl >= wmin && wl < lim) ? p : 0ull;
    }
  } else {
    // running group: gl = last staged group starting at or before word wl (-1: none)
    int wl = tid * GWPT - 1;
    const int num = 32 * (wl - D) + 31 - r0;
    int gl = num < 0 ? -1 : (int)((unsigned)num / (unsigned)gb);
    if (gl > ngl - 1) gl = ngl - 1;
    int Wt[GCOV];  // Wt[t]: first block word of group gl - t
#pragma unroll
    for (int t = 0; t < GCOV; t++) Wt[t] = ((r0 + (gl - t) * gb) >> 5) + D;
    int wnext = ((r0 + (gl + 1) * gb) >> 5) + D;  // first word of group gl + 1
#pragma unroll
    for (int k = 0; k <= GWPT; k++, wl++) {
      if (k > 0 && gl + 1 < ngl && wnext <= wl) {
        gl++;
#pragma unroll
        for (int t = GCOV - 1; t > 0; t--) Wt[t] = Wt[t - 1];
        Wt[0] = wnext;
        wnext = ((r0 + (gl + 1) * gb) >> 5) + D;
      }
      // branch-free: out-of-range terms read slot 0 and count as 0
      unsigned long long p = 0;
#pragma unroll
      for (int t = 0; t < GCOV; t++) {
        const int off = wl - Wt[t];
        const bool ok = gl - t >= 0 && off < a.gK;
        const unsigned i = ok ? (unsigned)((gl - t) * a.gKs + off) : 0u;
        const unsigned x = gl_w[i + (i >> 5)];
        p += ok ? x : 0u;
      }
      P[k] = (wl >= wmin && wl < lim) ? p : 0ull;
    }
  }
  if (a.add) {  // the addend's words, where they exist
#pragma unroll
    for (int k = 0; k <= GWPT; k++) {
      const long long w = wfirst - 1 + k;
      if (w >= 0 && w < (long long)a.CL && w < (long long)a.addl) P[k] += a.add[w];
    }
  }
  if (wfirst == 0) P[0] = 0;
  unsigned long long v[GWPT];
#pragma unroll
  for (int k = 0; k < GWPT; k++) v[k] = (P[k + 1] & 0xffffffffull) + (P[k] >> 32);
  unsigned f0 = 0, f1 = 1;
#pragma unroll
  for (int k = 0; k < GWPT; k++) {
    f0 = (unsigned)((v[k] + f0) >> 32);
    f1 = (unsigned)((v[k] + f1) >> 32);
  }
  const unsigned f = f0 | (f1 << 1);
  const int lane = tid & 63, wave = tid >> 6;
  unsigned inc = f;
  for (int off = 1; off < 64; off <<= 1) {
    unsigned gq = __shfl_up(inc, off, 64);
    if (lane >= off) inc = carry_compose(gq, inc);
  }
  unsigned excl = __shfl_up(inc, 1, 64);
  if (lane == 0) excl = 2u;
  if (lane == 63) wave_fn[wave] = inc;
  __syncthreads();
  unsigned pre = 2u;
  for (int q = 0; q < wave; q++) pre = carry_compose(pre, wave_fn[q]);
  if (tid == 0) {
    unsigned tot = 2u;
    for (int q = 0; q < GCT / 64; q++) tot = carry_compose(tot, wave_fn[q]);
    a.blk_fn[blockIdx.x] = (unsigned char)tot;
  }
  unsigned cin = carry_compose(pre, excl) & 1u;
  uint32_t o[GWPT];
#pragma unroll
  for (int k = 0; k < GWPT; k++) {
    unsigned long long sum = v[k] + cin;
    o[k] = (uint32_t)sum;
    cin = (unsigned)(sum >> 32);
  }
  if (wfirst + GWPT <= (long long)a.CL && ((reinterpret_cast<uintptr_t>(a.out + wfirst) & 15) == 0)) {
    uint4 *dst = reinterpret_cast<uint4 *>(a.out + wfirst);
#pragma unroll
    for (int q = 0; q < GWPT / 4; q++) dst[q] = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
  } else {
#pragma unroll
    for (int k = 0; k < GWPT; k++)
      if (wfirst + k < (long long)a.CL) a.out[wfirst + k] = o[k];
  }
}

// One workgroup per product: scan the block carry functions, then add the
// carry into every block whose carry-in is 1 (ripple through 0xffffffff words
// inside the block). Thread t owns blocks [t per, t per + per) with per a
// multiple of 16, so its function bytes arrive as 16-byte loads; the scan
// across threads uses wave shuffles; the first words of the carried blocks
// of a 16-block chunk are read together and take the +1 when it stops there;
// the other carried blocks (all-ones blocks, as in products close to a power
// of two, and +1s that run on) are marked for k_carry_apply.
static __global__ __launch_bounds__(1024) void k_carry_fix(CarryArgs a) {
  __shared__ unsigned wfn[16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  bind_carry(a, blockIdx.x);
  const unsigned long long nb = a.nblk;
  const unsigned long long per = (((nb + 1023) / 1024) + 15) & ~15ull;
  const unsigned long long b0 = (unsigned long long)tid * per;
  unsigned long long b1 = b0 + per;
  if (b1 > nb) b1 = nb;
  unsigned fn = 2u;  // identity
  for (unsigned long long c0 = b0; c0 < b1; c0 += 16) {
    const uint4 v = *reinterpret_cast<const uint4 *>(a.blk_fn + c0);
    const unsigned wv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 16; k++)
      if (c0 + k < b1) fn = carry_compose(fn, (wv[k >> 2] >> (8 * (k & 3))) & 3u);
  }
  unsigned inc = fn;
  for (int off = 1; off < 64; off <<= 1) {
    unsigned gq = __shfl_up(inc, off, 64);
    if (lane >= off) inc = carry_compose(gq, inc);
  }
  unsigned excl = __shfl_up(inc, 1, 64);
  if (lane == 0) excl = 2u;
  if (lane == 63) wfn[wave] = inc;
  __syncthreads();
  unsigned pre = 2u;
  for (int w = 0; w < wave; w++) pre = carry_compose(pre, wfn[w]);
  unsigned c = carry_compose(pre, excl) & 1u;
  const unsigned long long wpb = a.wpb;
  for (unsigned long long c0 = b0; c0 < b1; c0 += 16) {
    const uint4 v = *reinterpret_cast<const uint4 *>(a.blk_fn + c0);
    const unsigned wv[4] = {v.x, v.y, v.z, v.w};
    unsigned carried = 0, heavy = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      if (c0 + k < b1) {
        const unsigned f = (wv[k >> 2] >> (8 * (k & 3))) & 3u;
        if (c) {
          // a carried block whose function propagates holds only 0xffffffff
          // words (it becomes zero): left to k_carry_apply
          if (f == 2u)
            heavy |= 1u << k;
          else
            carried |= 1u << k;
        }
        c = (f >> c) & 1u;
      }
    }
    uint32_t x[16];
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = ((carried >> k) & 1u) ? a.out[(c0 + k) * wpb] : 0u;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      if (!((carried >> k) & 1u)) continue;
      if (x[k] != 0xffffffffu)
        a.out[(c0 + k) * wpb] = x[k] + 1u;
      else
        heavy |= 1u << k;  // the +1 runs on through 0xffffffff words
    }
    if (heavy)
      for (int k = 0; k < 16; k++)
        if ((heavy >> k) & 1u) a.blk_fn[c0 + k] |= 4u;
  }
}

// The carried blocks k_carry_fix marked (bit 2 of the block function): one
// workgroup per block adds the carry with all its threads, clearing a block of
// 0xffffffff words, or the 0xffffffff words below the first other word, which
// takes the +1 (a block whose function does not propagate has such a word).
static __global__ __launch_bounds__(256) void k_carry_apply(CarryArgs a) {
  __shared__ unsigned long long first;
  const int tid = threadIdx.x;
  bind_carry(a, blockIdx.y);
  if (blockIdx.x >= a.nblk) return;
  const unsigned f = a.blk_fn[blockIdx.x];
  if (!(f & 4u)) return;
  const unsigned long long w0 = (unsigned long long)blockIdx.x * a.wpb;
  const unsigned long long w1 = w0 + a.wpb < a.CL ? w0 + a.wpb : a.CL;
  if ((f & 3u) == 2u) {
    for (unsigned long long w = w0 + tid; w < w1; w += 256) a.out[w] = 0u;
    return;
  }
  if (tid == 0) first = w1;
  __syncthreads();
  for (unsigned long long w = w0 + tid; w < w1; w += 256)
    if (a.out[w] != 0xffffffffu) {
      atomicMin(&first, w);
      break;
    }
  __syncthreads();
  const unsigned long long fw = first;
  for (unsigned long long w = w0 + tid; w < fw; w += 256) a.out[w] = 0u;
  if (tid == 0 && fw < w1) a.out[fw] += 1u;
}

// --------------------------------------------------------------- utilities

static __global__ void k_zero_u32(unsigned int *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

static __global__ void k_fill_u32(uint32_t *p, unsigned long long n, uint32_t v) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long st = (unsigned long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = v;
}

}  // namespace bcg

namespace bcg {

// ------------------------------------------- one-workgroup products (batched)


// P_w for one word from coefficients held in LDS (see word_partials).
__device__ __forceinline__ unsigned long long word_partial_lds(const long long *coef, long long ncoef, int bits,
                                                               int cbits, long long w) {
  if (w < 0) return 0ull;
  long long lo_bit = 32 * w - cbits + 1;
  long long jlo = lo_bit <= 0 ? 0 : (lo_bit + bits - 1) / bits;
  long long jhi = (32 * w + 31) / bits;
  if (jhi >= ncoef) jhi = ncoef - 1;
  unsigned long long P = 0;
  for (long long j = jlo; j <= jhi; j++) {
    unsigned long long c = (unsigned long long)coef[j];
    long long sh = (long long)bits * j - 32 * w;
    if (sh >= 0)
      P += (unsigned long long)(uint32_t)(c << sh);
    else if (-sh < 64)
      P += (unsigned long long)(uint32_t)(c >> (-sh));
  }
  return P;
}

// Words out[0..CL) = (sum_j coef[j] 2^{bits j}) mod 2^{32 CL}, coefficients in
// LDS, carries resolved with a workgroup scan of carry functions; 64-bit word
// indices, every word's partial computed twice (for output lengths of 2^19
// words and more: a short product with a long addend).
template <int T>
__device__ void wg_carry_wide(const long long *coef, long long ncoef, int bits, int cbits, uint32_t *out, long long CL,
                         const uint32_t *add, long long addl, int tid) {
  __shared__ unsigned wave_fn[T / 64 + 1];
  const long long per = (CL + T - 1) / T;
  const long long w0 = (long long)tid * per;
  long long w1 = w0 + per;
  if (w1 > CL) w1 = CL;
  // this thread's carry function over words [w0, w1)
  unsigned f = 2u;  // identity
  {
    unsigned long long prev = word_partial_lds(coef, ncoef, bits, cbits, w0 - 1);
    if (add && w0 >= 1 && w0 - 1 < addl) prev += add[w0 - 1];
    unsigned c0 = 0, c1 = 1;
    for (long long w = w0; w < w1; w++) {
      unsigned long long P = word_partial_lds(coef, ncoef, bits, cbits, w);
      if (add && w < addl) P += add[w];
      unsigned long long v = (P & 0xffffffffull) + (prev >> 32);
      c0 = (unsigned)((v + c0) >> 32);
      c1 = (unsigned)((v + c1) >> 32);
      prev = P;
    }
    if (w0 < w1) f = c0 | (c1 << 1);
  }
  // inclusive scan within the wave
  const int lane = tid & 63, wave = tid >> 6;
  unsigned inc = f;
  for (int off = 1; off < 64; off <<= 1) {
    unsigned g = __shfl_up(inc, off, 64);
    if (lane >= off) inc = carry_compose(g, inc);
  }
  unsigned excl = __shfl_up(inc, 1, 64);
  if (lane == 0) excl = 2u;
  if (lane == 63) wave_fn[wave] = inc;
  __syncthreads();
  unsigned pre = 2u;
  for (int w = 0; w < wave; w++) pre = carry_compose(pre, wave_fn[w]);
  unsigned cin = carry_compose(pre, excl) & 1u;  // carry into word w0 (block carry-in is 0)
  // write the words
  unsigned long long prev = word_partial_lds(coef, ncoef, bits, cbits, w0 - 1);
  if (add && w0 >= 1 && w0 - 1 < addl) prev += add[w0 - 1];
  for (long long w = w0; w < w1; w++) {
    unsigned long long P = word_partial_lds(coef, ncoef, bits, cbits, w);
    if (add && w < addl) P += add[w];
    unsigned long long v = (P & 0xffffffffull) + (prev >> 32) + cin;
    out[w] = (uint32_t)v;
    cin = (unsigned)(v >> 32);
    prev = P;
  }
}

// P_w for word w < 2^19 from the coefficients in LDS, 32-bit index arithmetic:
// coefficient j (below 2^cbits) sits at bit b j, so words take j in
// [ceil((32 w - cbits + 1) / b), floor((32 w + 31) / b)].
__device__ __forceinline__ unsigned long long word_partial32(const long long *coef, int ncoef, int bits, int cbits,
                                                             float inv_b, int w) {
  if (w < 0) return 0ull;
  const int lo_bit = 32 * w - cbits + 1;
  const int jlo = lo_bit <= 0 ? 0 : fdiv_small(lo_bit + bits - 1, bits, inv_b);
  int jhi = fdiv_small(32 * w + 31, bits, inv_b);
  if (jhi >= ncoef) jhi = ncoef - 1;
  unsigned long long P = 0;
  int sh = bits * jlo - 32 * w;  // in (-cbits, 32)
  for (int j = jlo; j <= jhi; j++, sh += bits) {
    const unsigned long long c = (unsigned long long)coef[j];
    P += (unsigned long long)(uint32_t)(sh >= 0 ? c << sh : c >> -sh);
  }
  return P;
}

// Words out[0..CL) = (sum_j coef[j] 2^{bits j} + add) mod 2^{32 CL}, coefficients
// in LDS. Each thread takes CL / T consecutive words and writes them as for a
// carry-in of 0, computing every word's partial once, while it tracks its
// carry outs for carry-ins 0 and 1; a workgroup scan of those carry functions
// gives each thread its carry-in, and a carry-in of 1 then ripples through the
// thread's leading all-ones words.
template <int T>
__device__ void wg_carry(const long long *coef, long long ncoef, int bits, int cbits, uint32_t *out, long long CL,
                         const uint32_t *add, long long addl, int tid) {
  if (CL >= (1ll << 19) || ncoef >= (1ll << 30)) {  // 32 w + 31 < 2^24 below
    wg_carry_wide<T>(coef, ncoef, bits, cbits, out, CL, add, addl, tid);
    return;
  }
  __shared__ unsigned wave_fn[T / 64 + 1];
  const int cl = (int)CL, nc = (int)ncoef;
  const float inv_b = 1.0f / (float)bits;
  const int per = (cl + T - 1) / T;
  const int w0 = tid * per;
  const int w1 = w0 + per < cl ? w0 + per : cl;
  unsigned f = 2u;  // identity
  {
    unsigned long long prev = word_partial32(coef, nc, bits, cbits, inv_b, w0 - 1);
    if (add && w0 >= 1 && w0 - 1 < addl) prev += add[w0 - 1];
    unsigned c0 = 0, c1 = 1;
    for (int w = w0; w < w1; w++) {
      unsigned long long P = word_partial32(coef, nc, bits, cbits, inv_b, w);
      if (add && w < addl) P += add[w];
      const unsigned long long v = (P & 0xffffffffull) + (prev >> 32);
      const unsigned long long s0 = v + c0;
      out[w] = (uint32_t)s0;
      c0 = (unsigned)(s0 >> 32);
      c1 = (unsigned)((v + c1) >> 32);
      prev = P;
    }
    if (w0 < w1) f = c0 | (c1 << 1);
  }
  // inclusive scan within the wave
  const int lane = tid & 63, wave = tid >> 6;
  unsigned inc = f;
  for (int off = 1; off < 64; off <<= 1) {
    unsigned g = __shfl_up(inc, off, 64);
    if (lane >= off) inc = carry_compose(g, inc);
  }
  unsigned excl = __shfl_up(inc, 1, 64);
  if (lane == 0) excl = 2u;
  if (lane == 63) wave_fn[wave] = inc;
  __syncthreads();
  unsigned pre = 2u;
  for (int w = 0; w < wave; w++) pre = carry_compose(pre, wave_fn[w]);
  const unsigned cin = carry_compose(pre, excl) & 1u;  // carry into word w0 (block carry-in is 0)
  if (cin) {
    for (int w = w0; w < w1; w++) {
      const uint32_t x = out[w] + 1u;
      out[w] = x;
      if (x) break;
    }
  }
}

// All products of one transform length class L (one-pass plans: N = L).
struct ClassArgs {
  const MulDesc *desc;
  int count;
  Roots roots;
  const cd *tl;
  unsigned int *margins;  // per product, float bits of the worst rounding distance
  int *next;              // work counter, zero at launch
};

// Persistent workgroups pull products of one length class from a counter and
// run each product whole: pack, forward, pointwise, inverse, rounding and the
// carry resolve, with the coefficients never leaving LDS.
template <int L>
__global__ __launch_bounds__((Tile<L, 1>::T)) void k_mul1(ClassArgs ca) {
  using TC = Tile<L, 1>;
  __shared__ cd lds[TC::LDS_ELEMS];
  __shared__ int job;
  const int tid = threadIdx.x;
  constexpr int R0 = 1 << TC::FIRSTB;
  constexpr int PER0 = TC::EPT / R0;
  constexpr int SUB0 = L / R0;
  for (;;) {
    if (tid == 0) job = atomicAdd(ca.next, 1);
    __syncthreads();
    const int idx = job;
    __syncthreads();
    if (idx >= ca.count) break;
    const MulDesc d = ca.desc[idx];
    PassArgs a;
    a.data = nullptr;
    a.data_b = nullptr;
    a.wa = d.a;
    a.wb = d.b;
    a.al = d.al;
    a.bl = d.bl;
    a.coef = nullptr;
    a.N = L;
    a.M = 1;
    a.S = L;
    a.tw_mul = 4;
    a.tl = ca.tl;
    a.roots = ca.roots;
    a.bits = d.bits;
    a.ntiles = 1;
    a.margin = nullptr;
    a.inv_n = 1.0 / (double)L;
    cd xa[TC::EPT];
    if (d.a == d.b && d.al == d.bl)
      mid_pipeline<L, true, true>(a, lds, 0, xa, tid);
    else
      mid_pipeline<L, false, true>(a, lds, 0, xa, tid);
    __syncthreads();  // LDS now becomes the coefficient array
    long long *coef = (long long *)lds;
    double worst = 0.0;
#pragma unroll
    for (int i = 0; i < PER0; i++) {
      int q = tid + i * TC::T;
#pragma unroll
      for (int t = 0; t < R0; t++) {
        int j = q + t * SUB0;
        cd f = cmulc(xa[i * R0 + t], omega_pow(a.roots, (unsigned long long)j));
        const bool fast = d.cbits <= 51;
        coef[j] = round_coef(f.x * a.inv_n, fast, worst);
        coef[j + L] = round_coef(f.y * a.inv_n, fast, worst);
      }
    }
    margin_reduce(ca.margins + idx, worst);
    __syncthreads();
    wg_carry<TC::T>(coef, 2 * L, d.bits, d.cbits, d.c, (long long)d.cl, d.d, (long long)d.dl, tid);
    __syncthreads();
  }
}

}  // namespace bcg

namespace bcg {

// Batched schoolbook products (CL < 100 words, AL + BL <= 128): one wave per
// product; lane w sums column w of 32x32-bit partial products (low and high
// halves apart), lane 0 then runs the carry chain, adding the optional addend.
static __global__ __launch_bounds__(64) void k_basecase_batch(const MulDesc *desc, int count) {
  __shared__ unsigned long long lo[128], hi[128];
  const int p = blockIdx.x;
  if (p >= count) return;
  const MulDesc d = desc[p];
  const int W = (int)(d.al + d.bl);
  for (int w = threadIdx.x; w < W; w += 64) {
    unsigned long long l = 0, h = 0;
    int i0 = w >= (int)d.bl ? w - (int)d.bl + 1 : 0;
    for (int i = i0; i < (int)d.al && i <= w; i++) {
      unsigned long long pr = (unsigned long long)d.a[i] * d.b[w - i];
      l += pr & 0xffffffffull;
      h += pr >> 32;
    }
    lo[w] = l;
    hi[w] = h;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long carry = 0;
    for (unsigned long long w = 0; w < d.cl; w++) {
      unsigned long long v = carry;
      if (w < (unsigned long long)W) v += lo[w];
      if (w >= 1 && w - 1 < (unsigned long long)W) v += hi[w - 1];
      if (d.d && w < d.dl) v += d.d[w];
      d.c[w] = (uint32_t)v;
      carry = v >> 32;
    }
  }
}

}  // namespace bcg

namespace bcg {

}  // namespace bcg
