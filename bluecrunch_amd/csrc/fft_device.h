// Device code of the FP64 FFT multiply for gfx950 (MI355X).
//
// Transform structure (all kernels):
//  * A length-N transform (N = 2^k or 3*2^k) is split into P <= 3 HBM passes
//    N = L_1 * ... * L_P. Pass p views its span S_p = L_p * M_p
//    (M_p = L_{p+1}...L_P) as L_p rows of M_p columns and runs length-L_p DFTs
//    down the columns, C adjacent columns per workgroup so every global
//    access is a C*16-byte contiguous run.
//  * Inside a workgroup the length-L DFT is an in-place radix-2 decimation in
//    frequency held in LDS; groups of up to 4 radix-2 levels (radix 16) run in
//    registers between LDS exchanges. Output rows come out bit-reversed, so a
//    whole transform lands in exactly the order of the reference fft_forward
//    (fft/fft.c:43-96: bit-reversed, radix-3 digit innermost, e^{+2 pi i/N}).
//  * The inverse is the mirror image: radix-2 decimation in time with
//    conjugate twiddles (fft/fft.c:289-329), bit-reversed input, natural output.
//
// The multiply (bigfloat/mul.c:108-174 semantics: exact product) uses a
// right-angle (negacyclic) convolution: 2N real b-bit digits become N complex
// points u_j = (a_j + i a_{j+N}) psi^j with psi = e^{i pi / 2N}, so no zero
// padding and no real/complex untangling is needed. After the inverse,
// (c_j + i c_{j+N}) = w_j psi^{-j} / N.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bcg {

typedef double2 cd;

__device__ __forceinline__ cd cadd(cd a, cd b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cd csub(cd a, cd b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cd cmul(cd a, cd b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ cd cmulc(cd a, cd b) {
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// Streaming (nontemporal) 16-byte loads and stores for the transform passes:
// every pass touches each point once and the next pass comes back to it only
// after gigabytes of other traffic, so the points are not kept in the caches
// (measured 6 % faster over the whole multiply at 2^24 words).
typedef double nt_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ cd ld_nt(const cd *p) {
  const nt_d2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_d2 *>(p));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st_nt(cd *p, cd v) {
  nt_d2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<nt_d2 *>(p));
}
// Cache policy of a pass: ordinary loads and stores where the next launch
// reads the points back while they are still in the memory-side cache (the
// chunked middle passes of a three-pass product), streaming ones otherwise.
// CACHE bit 0: ordinary loads; bit 1: ordinary stores.
template <bool PLAIN>
__device__ __forceinline__ cd ld_pol(const cd *p) {
  if constexpr (PLAIN) return *p;
  else return ld_nt(p);
}
template <bool PLAIN>
__device__ __forceinline__ void st_pol(cd *p, cd v) {
  if constexpr (PLAIN) *p = v;
  else st_nt(p, v);
}
// e^{+2 pi i j / 16}, j = 0..15, to 20 significant digits.
#define BCG_C1 0.92387953251128675613
#define BCG_R2 0.70710678118654752440
__device__ __forceinline__ cd w16(int j) {
  // j is a compile-time constant at every call site (fully unrolled loops).
  const double c[16] = {1.0, BCG_C1, BCG_R2, 0.38268343236508977173, 0.0, -0.38268343236508977173, -BCG_R2, -BCG_C1,
                        -1.0, -BCG_C1, -BCG_R2, -0.38268343236508977173, 0.0, 0.38268343236508977173, BCG_R2, BCG_C1};
  const double s[16] = {0.0, 0.38268343236508977173, BCG_R2, BCG_C1, 1.0, BCG_C1, BCG_R2, 0.38268343236508977173,
                        0.0, -0.38268343236508977173, -BCG_R2, -BCG_C1, -1.0, -BCG_C1, -BCG_R2, -0.38268343236508977173};
  return make_double2(c[j & 15], s[j & 15]);
}

constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }

// Length descriptor: L = 2^LOG or 3 * 2^LOG.
template <int L>
struct Len {
  static constexpr bool THREE = (L % 3) == 0;
  static constexpr int P2 = THREE ? L / 3 : L;
  static constexpr int LOG = ilog2c(P2);
};

// Per-(L, C) tile configuration: E elements, EPT per thread, T threads and the
// radix-2^b stage list (first stage takes the remainder bits).
template <int L, int C>
struct Tile {
  static constexpr int E = L * C;
  static constexpr bool THREE = Len<L>::THREE;
  static constexpr int LOG = Len<L>::LOG;
  static constexpr int EPT = THREE ? 24 : 16;
  static constexpr int T = E / EPT;
  static constexpr int MAXB = THREE ? 3 : 4;
  static constexpr int NST2 = (LOG + MAXB - 1) / MAXB;  // radix-2^b stages
  static constexpr int FIRSTB = LOG - MAXB * (NST2 - 1);
  // LDS index with one pad element every 16 to spread power-of-two strides.
  static constexpr int LDS_ELEMS = E + E / 16;
  static_assert(E % EPT == 0, "tile must split evenly");
  static_assert(T >= 1 && T <= 1024, "bad thread count");
};

__device__ __forceinline__ int pad16(int e) { return e + (e >> 4); }

// pad16(row * C + c) for a column c < C, written linear in row when 16 | C, so
// the per-row terms of an unrolled loop become instruction offsets instead of
// one address register per row.
template <int C>
__device__ __forceinline__ int slot(int row, int c) {
  if constexpr (C % 16 == 0) return row * (C + C / 16) + c + (c >> 4);
  else return pad16(row * C + c);
}

// pad16(q + t * S), linear in t when 16 | S (q >= 0).
template <int S>
__device__ __forceinline__ int slot_step(int q, int t) {
  if constexpr (S % 16 == 0) return pad16(q) + t * (S + S / 16);
  else return pad16(q + t * S);
}

__device__ __forceinline__ unsigned brev(unsigned v, int bits) {
  return bits ? (__builtin_bitreverse32(v) >> (32 - bits)) : 0u;
}

// Frequency index held at row r after an in-place DIF of length L
// (bit-reversed 2-part, radix-3 digit innermost).
template <int L>
__device__ __forceinline__ unsigned freq_of_row(unsigned r) {
  if (Len<L>::THREE) return brev(r / 3u, Len<L>::LOG) + (r % 3u) * (unsigned)Len<L>::P2;
  return brev(r, Len<L>::LOG);
}

// Root-of-unity tables for one transform length N: Omega = e^{2 pi i / 4N}.
//   Omega^u = hi[u >> sh] * lo[u & (2^sh - 1)],  u in [0, 4N)
//   tl_L[t] = e^{2 pi i t / L} for t < L/2, one table per pass length L.
//   Without tables (omega_direct): 4N = 8 q8 with q8 = 2^qa or 3 * 2^qa and
//   step = (pi/4) / q8 rounded to double.
struct Roots {
  const cd *lo;
  const cd *hi;
  int sh;
  unsigned long long four_n;
  unsigned long long q8;
  int qa;
  int q3;
  double step;
};

__device__ __forceinline__ cd omega_pow(const Roots &r, unsigned long long u) {
  return cmul(r.hi[u >> r.sh], r.lo[u & ((1ull << r.sh) - 1)]);
}

// Omega^u for 0 <= u < 4N, computed in registers. The octant o = floor(u / q8)
// is split off exactly in integers; the rest is an angle y in [0, pi/4]
// (measured back from the next octant boundary when o is odd) whose cosine and
// sine come from their Taylor series through y^16 and y^17 (truncation below
// 3e-18 at y = pi/4); octant symmetries then place the point. No memory access,
// so twiddles of strided passes cost ALU work instead of scattered loads.
__host__ __device__ __forceinline__ cd omega_direct(const Roots &r, unsigned long long u) {
  const unsigned x = (unsigned)(u >> r.qa);             // < 8, or < 24 when q8 = 3 * 2^qa
  const unsigned o = r.q3 ? (x * 11u) >> 5 : x;         // floor(x / 3) for x < 24
  unsigned long long rem = u - (unsigned long long)o * r.q8;
  if (o & 1) rem = r.q8 - rem;
  const double y = (double)rem * r.step;
  const double t = y * y;
  double ps = 1.0 / 355687428096000.0;                  // 1/17!
  ps = fma(ps, t, -1.0 / 1307674368000.0);
  ps = fma(ps, t, 1.0 / 6227020800.0);
  ps = fma(ps, t, -1.0 / 39916800.0);
  ps = fma(ps, t, 1.0 / 362880.0);
  ps = fma(ps, t, -1.0 / 5040.0);
  ps = fma(ps, t, 1.0 / 120.0);
  ps = fma(ps, t, -1.0 / 6.0);
  const double sn = fma(y * t, ps, y);
  double pc = 1.0 / 20922789888000.0;                   // 1/16!
  pc = fma(pc, t, -1.0 / 87178291200.0);
  pc = fma(pc, t, 1.0 / 479001600.0);
  pc = fma(pc, t, -1.0 / 3628800.0);
  pc = fma(pc, t, 1.0 / 40320.0);
  pc = fma(pc, t, -1.0 / 720.0);
  pc = fma(pc, t, 1.0 / 24.0);
  pc = fma(pc, t, -0.5);
  const double cs = fma(t, pc, 1.0);
  // o = 0..7: (c,s) (s,c) (-s,c) (-c,s) (-c,-s) (-s,-c) (s,-c) (c,-s)
  const bool swap = ((o + 1) & 2) != 0;
  double re = swap ? sn : cs, im = swap ? cs : sn;
  if ((o + 2) & 4) re = -re;
  if (o & 4) im = -im;
  return make_double2(re, im);
}

// e^{2 pi i h / 64}, h = 0..15 (compile-time h at every call site).
__device__ __forceinline__ cd w64(int h) {
  const double c[16] = {1.0,
                        0.995184726672196886245,
                        0.980785280403230449126,
                        0.956940335732208864936,
                        0.923879532511286756128,
                        0.881921264348355029713,
                        0.831469612302545237079,
                        0.773010453362736960811,
                        0.707106781186547524401,
                        0.634393284163645498215,
                        0.555570233019602224743,
                        0.471396736825997648556,
                        0.382683432365089771728,
                        0.290284677254462367636,
                        0.195090322016128267848,
                        0.098017140329560601994};
  return make_double2(c[h & 15], c[(16 - h) & 15] * (h ? 1.0 : 0.0));
}

// Sixteen twiddles base * g^j with j = brev4(i) for i = 0, 1, ..., 15 (the
// frequency order of a thread's rows): base, g, g^2, g^4 and g^8 come from
// omega_direct, the other eleven are products along the bits of j, at most
// four deep. at(i) must be called with i = 0..15 in increasing order.
struct Walk16 {
  cd base, g1, g2, g4, g8, p3, p2, p1;
  __device__ __forceinline__ void init(const Roots &r, unsigned long long u0, unsigned long long du) {
    base = omega_direct(r, u0);
    g1 = omega_direct(r, du);
    g2 = omega_direct(r, 2 * du);
    g4 = omega_direct(r, 4 * du);
    g8 = omega_direct(r, 8 * du);
  }
  // Same, when lanes l, l + C, l + 2C, l + 3C of each 4C-lane block share du
  // (same column): each computes one of g, g^2, g^4, g^8 and the four are
  // exchanged by lane shuffles (4C <= 64, all 4C lanes active; else init).
  template <int C>
  __device__ __forceinline__ void init_shared(const Roots &r, unsigned long long u0, unsigned long long du) {
    if constexpr (4 * C > 64) {  // wider tiles: every lane computes its own
      init(r, u0, du);
      return;
    }
    const int lane = threadIdx.x & 63;
    const int q = (lane / C) & 3;
    const int src0 = (lane % C) + (lane / (4 * C)) * (4 * C);
    base = omega_direct(r, u0);
    const cd mine = omega_direct(r, du << q);
    g1 = make_double2(__shfl(mine.x, src0, 64), __shfl(mine.y, src0, 64));
    g2 = make_double2(__shfl(mine.x, src0 + C, 64), __shfl(mine.y, src0 + C, 64));
    g4 = make_double2(__shfl(mine.x, src0 + 2 * C, 64), __shfl(mine.y, src0 + 2 * C, 64));
    g8 = make_double2(__shfl(mine.x, src0 + 3 * C, 64), __shfl(mine.y, src0 + 3 * C, 64));
  }
  __device__ __forceinline__ cd at(int i) {
    if ((i & 7) == 0) p3 = (i & 8) ? cmul(base, g1) : base;
    if ((i & 3) == 0) p2 = (i & 4) ? cmul(p3, g2) : p3;
    if ((i & 1) == 0) p1 = (i & 2) ? cmul(p2, g4) : p2;
    return (i & 1) ? cmul(p1, g8) : p1;
  }
};

// ---------------------------------------------------------------- butterflies

// Level twiddle bases of dif_levels / dit_levels: base[l] = tl[q (tstride << l)]
// (loaded ahead of the data where a kernel wants their latency hidden).
template <int RB>
__device__ __forceinline__ void level_bases(const cd *__restrict__ tl, int q, int tstride, cd (&base)[RB]) {
#pragma unroll
  for (int l = 0; l < RB; l++) base[l] = tl[q * (tstride << l)];
}

// RB radix-2 DIF levels on R = 2^RB registers holding rows q + t*(span/R).
// Level l pairs t, t+half (half = R >> (l+1)); twiddle e^{2 pi i n / (span/2^l)}
// with n = q + (t mod half)*(span/R) = base_l * w_{2 half}^{t mod half}.
template <int RB>
__device__ __forceinline__ void dif_levels(cd (&x)[1 << RB], const cd (&tb)[RB]) {
  constexpr int R = 1 << RB;
#pragma unroll
  for (int l = 0; l < RB; l++) {
    const int half = R >> (l + 1);
    const cd base = tb[l];
#pragma unroll
    for (int t = 0; t < R; t++) {
      if ((t & (2 * half - 1)) < half) {
        const int j = t & (2 * half - 1);
        cd a = x[t], b = x[t + half];
        x[t] = cadd(a, b);
        cd d = csub(a, b);
        cd w = (j == 0) ? base : cmul(base, w16(j * (16 / (2 * half))));
        x[t + half] = cmul(d, w);
      }
    }
  }
}

// The same with each level's base loaded where the level starts.
template <int RB>
__device__ __forceinline__ void dif_levels(cd (&x)[1 << RB], const cd *__restrict__ tl, int q, int tstride) {
  constexpr int R = 1 << RB;
#pragma unroll
  for (int l = 0; l < RB; l++) {
    const int half = R >> (l + 1);
    const cd base = tl[q * (tstride << l)];
#pragma unroll
    for (int t = 0; t < R; t++) {
      if ((t & (2 * half - 1)) < half) {
        const int j = t & (2 * half - 1);
        cd a = x[t], b = x[t + half];
        x[t] = cadd(a, b);
        cd d = csub(a, b);
        cd w = (j == 0) ? base : cmul(base, w16(j * (16 / (2 * half))));
        x[t + half] = cmul(d, w);
      }
    }
  }
}

// Mirror image: radix-2 DIT levels with conjugate twiddles (c = b conj(w);
// a' = a + c; b' = a - c), bit-reversed input to natural output.
template <int RB>
__device__ __forceinline__ void dit_levels(cd (&x)[1 << RB], const cd (&tb)[RB]) {
  constexpr int R = 1 << RB;
#pragma unroll
  for (int l = RB - 1; l >= 0; l--) {
    const int half = R >> (l + 1);
    const cd base = tb[l];
#pragma unroll
    for (int t = 0; t < R; t++) {
      if ((t & (2 * half - 1)) < half) {
        const int j = t & (2 * half - 1);
        cd a = x[t];
        cd w = (j == 0) ? base : cmul(base, w16(j * (16 / (2 * half))));
        cd c = cmulc(x[t + half], w);
        x[t] = cadd(a, c);
        x[t + half] = csub(a, c);
      }
    }
  }
}

template <int RB>
__device__ __forceinline__ void dit_levels(cd (&x)[1 << RB], const cd *__restrict__ tl, int q, int tstride) {
  constexpr int R = 1 << RB;
#pragma unroll
  for (int l = RB - 1; l >= 0; l--) {
    const int half = R >> (l + 1);
    const cd base = tl[q * (tstride << l)];
#pragma unroll
    for (int t = 0; t < R; t++) {
      if ((t & (2 * half - 1)) < half) {
        const int j = t & (2 * half - 1);
        cd a = x[t];
        cd w = (j == 0) ? base : cmul(base, w16(j * (16 / (2 * half))));
        cd c = cmulc(x[t + half], w);
        x[t] = cadd(a, c);
        x[t + half] = csub(a, c);
      }
    }
  }
}

// d e^{+2 pi i k / 16} and d e^{-2 pi i k / 16} for k in [0, 8), k a constant
// after unrolling: the factors 1 and i are exact swaps, (+-1 + i)/sqrt 2 one
// addition and one product per part, the rest a constant complex product.
__device__ __forceinline__ cd mul_w16(cd d, int k) {
  switch (k) {
    case 0: return d;
    case 2: return make_double2((d.x - d.y) * BCG_R2, (d.x + d.y) * BCG_R2);
    case 4: return make_double2(-d.y, d.x);
    case 6: return make_double2(-(d.x + d.y) * BCG_R2, (d.x - d.y) * BCG_R2);
    default: return cmul(d, w16(k));
  }
}
__device__ __forceinline__ cd mulc_w16(cd d, int k) {
  switch (k) {
    case 0: return d;
    case 2: return make_double2((d.x + d.y) * BCG_R2, (d.y - d.x) * BCG_R2);
    case 4: return make_double2(d.y, -d.x);
    case 6: return make_double2((d.y - d.x) * BCG_R2, -(d.x + d.y) * BCG_R2);
    default: return cmulc(d, w16(k));
  }
}

// dif_levels / dit_levels for q = 0 (the innermost stage of a tile): every
// twiddle is a sixteenth root of unity known at compile time.
template <int RB>
__device__ __forceinline__ void dif_levels_q0(cd (&x)[1 << RB]) {
  constexpr int R = 1 << RB;
#pragma unroll
  for (int l = 0; l < RB; l++) {
    const int half = R >> (l + 1);
#pragma unroll
    for (int t = 0; t < R; t++) {
      if ((t & (2 * half - 1)) < half) {
        const int j = t & (2 * half - 1);
        cd a = x[t], b = x[t + half];
        x[t] = cadd(a, b);
        x[t + half] = mul_w16(csub(a, b), j * (16 / (2 * half)));
      }
    }
  }
}

template <int RB>
__device__ __forceinline__ void dit_levels_q0(cd (&x)[1 << RB]) {
  constexpr int R = 1 << RB;
#pragma unroll
  for (int l = RB - 1; l >= 0; l--) {
    const int half = R >> (l + 1);
#pragma unroll
    for (int t = 0; t < R; t++) {
      if ((t & (2 * half - 1)) < half) {
        const int j = t & (2 * half - 1);
        cd a = x[t];
        cd c = mulc_w16(x[t + half], j * (16 / (2 * half)));
        x[t] = cadd(a, c);
        x[t + half] = csub(a, c);
      }
    }
  }
}

// 3-point DFT, natural order in and out, w = e^{+-2 pi i / 3}
// (include/intrinsic.h:270-363 dft_3p / dft_3p_inv).
template <bool INV>
__device__ __forceinline__ void dft3(cd &a, cd &b, cd &c) {
  const double wr = -0.5;
  const double wi = INV ? -0.86602540378443864676 : 0.86602540378443864676;
  cd s = cadd(b, c);
  cd d = csub(b, c);
  cd a0 = cadd(a, s);
  // a + s*wr  +-  i*wi*d
  cd m = make_double2(a.x + wr * s.x, a.y + wr * s.y);
  cd r = make_double2(-wi * d.y, wi * d.x);
  a = a0;
  b = cadd(m, r);
  c = csub(m, r);
}

// --------------------------------------------------------- LDS stage helpers

// One radix-2^RB stage of span `span` over the whole tile, LDS in and out.
template <int L, int C, int RB, bool INV>
__device__ __forceinline__ void lds_stage(cd *lds, int span, const cd *__restrict__ tl, int tid) {
  using TC = Tile<L, C>;
  constexpr int R = 1 << RB;
  constexpr int PER = TC::EPT / R;
  const int sub = span / R;
#pragma unroll
  for (int i = 0; i < PER; i++) {
    int bf = tid + i * TC::T;
    int c = bf % C;
    int rest = bf / C;
    int q = rest % sub;
    int blk = rest / sub;
    int row0 = blk * span + q;
    cd x[R];
#pragma unroll
    for (int t = 0; t < R; t++) x[t] = lds[slot<C>(row0 + t * sub, c)];
    if (INV)
      dit_levels<RB>(x, tl, q, L / span);
    else
      dif_levels<RB>(x, tl, q, L / span);
#pragma unroll
    for (int t = 0; t < R; t++) lds[slot<C>(row0 + t * sub, c)] = x[t];
  }
}

template <int L, int C, bool INV>
__device__ __forceinline__ void lds_radix3(cd *lds, int tid) {
  using TC = Tile<L, C>;
  constexpr int PER = TC::EPT / 3;
#pragma unroll
  for (int i = 0; i < PER; i++) {
    int bf = tid + i * TC::T;
    int c = bf % C;
    int blk = bf / C;
    const int r3 = blk * 3;
    cd a = lds[slot<C>(r3, c)], b = lds[slot<C>(r3 + 1, c)], d = lds[slot<C>(r3 + 2, c)];
    dft3<INV>(a, b, d);
    lds[slot<C>(r3, c)] = a;
    lds[slot<C>(r3 + 1, c)] = b;
    lds[slot<C>(r3 + 2, c)] = d;
  }
}

// Runtime dispatch on the radix bits of a stage (1..4).
template <int L, int C, bool INV>
__device__ __forceinline__ void lds_stage_any(cd *lds, int rb, int span, const cd *__restrict__ tl, int tid) {
  switch (rb) {
    case 1: lds_stage<L, C, 1, INV>(lds, span, tl, tid); break;
    case 2: lds_stage<L, C, 2, INV>(lds, span, tl, tid); break;
    case 3: lds_stage<L, C, 3, INV>(lds, span, tl, tid); break;
    default: lds_stage<L, C, 4, INV>(lds, span, tl, tid); break;
  }
}

// Stage bit widths: stage 0 takes FIRSTB bits, the rest MAXB.
template <int L, int C>
__device__ __forceinline__ int stage_bits(int s) {
  return s == 0 ? Tile<L, C>::FIRSTB : Tile<L, C>::MAXB;
}

// Forward stages 1..NST2-1 entirely in LDS (stage 0 has been done already),
// then the radix-3 stage when L = 3*2^k. Ends with a barrier.
template <int L, int C>
__device__ __forceinline__ void forward_rest(cd *lds, const cd *__restrict__ tl, int tid) {
  using TC = Tile<L, C>;
  int span = L >> TC::FIRSTB;
  for (int s = 1; s < TC::NST2; s++) {
    int rb = stage_bits<L, C>(s);
    lds_stage_any<L, C, false>(lds, rb, span, tl, tid);
    __syncthreads();
    span >>= rb;
  }
  if (TC::THREE) {
    lds_radix3<L, C, false>(lds, tid);
    __syncthreads();
  }
}

// Inverse stages from the innermost up to (not including) stage 0.
template <int L, int C>
__device__ __forceinline__ void inverse_front(cd *lds, const cd *__restrict__ tl, int tid) {
  using TC = Tile<L, C>;
  if (TC::THREE) {
    lds_radix3<L, C, true>(lds, tid);
    __syncthreads();
  }
  int span = TC::THREE ? 3 : 1;
  for (int s = TC::NST2 - 1; s >= 1; s--) {
    int rb = stage_bits<L, C>(s);
    span <<= rb;
    lds_stage_any<L, C, true>(lds, rb, span, tl, tid);
    __syncthreads();
  }
}

}  // namespace bcg
