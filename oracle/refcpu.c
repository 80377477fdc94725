/*
 * refcpu -- CPU restatement of the reference FFT bigfloat multiply.
 *
 * TEST INFRASTRUCTURE: the oracle and the CPU baseline. Never linked into the
 * product. Written from reading the reference's behaviour; no reference source
 * is copied. Cilk Plus spawn/for are replaced with OpenMP tasks/for at the same
 * thresholds (fft/fft.c:18 60000 points).
 *
 * Documented deviations from the reference at its HEAD snapshot:
 *  - fft_forward never runs the Bailey debug block (fft/fft.c:397-409), so
 *    power-of-two lengths > 30 are transformed once (the reference transforms
 *    twice or aborts there).
 *  - _fft_mul clamps bits per point below at 7 (bigfloat/mul.c:109-124 would
 *    select < 7 above CL = 4,000,000 words, which fft/convert.c:66-69 rejects).
 *  - int_to_fft / fft_to_int with AL == 0 touch nothing (the reference reads or
 *    writes word 0 there, fft/convert.c generated loops).
 *  - bits per point may be any value in [1, 30] (the reference generates 7..22).
 *  - bigfloat_mul computes the product out of place before it replaces the
 *    target mantissa, so the in-place realloc path keeps the truncation offset
 *    (bigfloat/mul.c:611-617 drops it; no reference caller reaches that path).
 */
#include "refcpu.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SPAWN_LENGTH 60000 /* fft/fft.c:18 */
#define MAX_TABLES 125     /* include/fft.h:5 */

typedef struct {
  double re, im;
} cplx;

/* ---------------------------------------------------------------- lengths */

/* fft/fft.c:22-41: smallest 2*2^k or 3*2^k >= n (never below 2). */
size_t ref_fft_length(size_t n) {
  size_t base = 1;
  for (;;) {
    if (2 * base >= n) return 2 * base;
    if (3 * base >= n) return 3 * base;
    base *= 2;
  }
}

/* fft/table.c:155-179 (release branch): slot of the table whose half length is h. */
size_t ref_table_select_raw(size_t h) {
  unsigned ones = (unsigned)__builtin_popcountl(h);
  size_t tz = (size_t)__builtin_ctzl(h);
  return ones == 1 ? 2 * (tz - 1) : 2 * tz + 1;
}

/* ----------------------------------------------------------------- tables */

static cplx *g_tables[MAX_TABLES];
static int g_table_size = 0;
static pthread_mutex_t g_table_lock = PTHREAD_MUTEX_INITIALIZER;

/* fft/table.c:189-209: L/2 twiddles e^{+2 pi i c / L} from libm cos/sin. */
static void build_table(size_t L) {
  size_t slot = ref_table_select_raw(L / 2);
  if (slot >= MAX_TABLES || g_tables[slot]) return;
  double step = 2 * M_PI / (double)L;
  cplx *t = (cplx *)malloc(sizeof(cplx) * (L / 2));
  for (size_t c = 0; c < L / 2; c++) {
    double ang = step * (double)c;
    t[c].re = cos(ang);
    t[c].im = sin(ang);
  }
  g_tables[slot] = t;
}

/* fft/table.c:211-228: tables for every 2^k and 3*2^k length up to `length`. */
void ref_ensure_table(size_t length) {
  pthread_mutex_lock(&g_table_lock);
  size_t k = 2;
  for (;;) {
    if (2 * k > length) break;
    build_table(2 * k);
    if (3 * k > length) break;
    build_table(3 * k);
    k *= 2;
  }
  int size = (int)ref_table_select_raw(k) + 1;
  if (size > g_table_size) g_table_size = size;
  pthread_mutex_unlock(&g_table_lock);
}

void ref_free_table(void) {
  pthread_mutex_lock(&g_table_lock);
  for (int i = 0; i < MAX_TABLES; i++) {
    free(g_tables[i]);
    g_tables[i] = NULL;
  }
  g_table_size = 0;
  pthread_mutex_unlock(&g_table_lock);
}

int ref_table_size(void) { return g_table_size; }

static const cplx *table_for(size_t L) { return g_tables[ref_table_select_raw(L / 2)]; }

/* -------------------------------------------------------------- small DFTs */

/* include/intrinsic.h:182-188 dft_2p (also the inverse leaf). */
static inline void dft2(cplx *v) {
  cplx a = v[0], b = v[1];
  v[0].re = a.re + b.re; v[0].im = a.im + b.im;
  v[1].re = a.re - b.re; v[1].im = a.im - b.im;
}

/* include/intrinsic.h:270-363 dft_3p / dft_3p_inv: natural order, w = e^{+-2 pi i/3}. */
static inline void dft3(cplx *v, double sgn) {
  const double wr = -0.5, wi = sgn * 0.86602540378443864676372317075293618;
  cplx a = v[0], b = v[1], c = v[2];
  /* b*w, c*w^2 (w^2 = conj w) */
  double bwr = b.re * wr - b.im * wi, bwi = b.im * wr + b.re * wi;
  double cwr = c.re * wr + c.im * wi, cwi = c.im * wr - c.re * wi;
  /* b*w^2, c*w^4 = c*w */
  double bvr = b.re * wr + b.im * wi, bvi = b.im * wr - b.re * wi;
  double cvr = c.re * wr - c.im * wi, cvi = c.im * wr + c.re * wi;
  v[0].re = a.re + b.re + c.re; v[0].im = a.im + b.im + c.im;
  v[1].re = a.re + bwr + cwr;   v[1].im = a.im + bwi + cwi;
  v[2].re = a.re + bvr + cvr;   v[2].im = a.im + bvi + cvi;
}

/* ------------------------------------------------------------- transforms */

/* fft/fft.c:43-96: radix-2 DIF, a' = a+b, b' = (a-b) w^n, then both halves;
 * output in bit-reversed order (radix-3 digit innermost for 3*2^k). */
static void fwd_rec(cplx *T, size_t L) {
  if (L == 2) { dft2(T); return; }
  if (L == 3) { dft3(T, 1.0); return; }
  size_t h = L / 2;
  const cplx *w = table_for(L);
  for (size_t n = 0; n < h; n++) {
    cplx a = T[n], b = T[n + h];
    double dr = a.re - b.re, di = a.im - b.im;
    T[n].re = a.re + b.re;
    T[n].im = a.im + b.im;
    T[n + h].re = dr * w[n].re - di * w[n].im;
    T[n + h].im = di * w[n].re + dr * w[n].im;
  }
  if (L >= SPAWN_LENGTH) {
#pragma omp task default(shared) if (L >= SPAWN_LENGTH)
    fwd_rec(T, h);
    fwd_rec(T + h, h);
#pragma omp taskwait
  } else {
    fwd_rec(T, h);
    fwd_rec(T + h, h);
  }
}

/* fft/fft.c:289-329: radix-2 DIT, halves first, then c = b conj(w^n),
 * a' = a+c, b' = a-c; bit-reversed input, natural unscaled output. */
static void inv_rec(cplx *T, size_t L) {
  if (L == 2) { dft2(T); return; }
  if (L == 3) { dft3(T, -1.0); return; }
  size_t h = L / 2;
  if (L >= SPAWN_LENGTH) {
#pragma omp task default(shared)
    inv_rec(T, h);
    inv_rec(T + h, h);
#pragma omp taskwait
  } else {
    inv_rec(T, h);
    inv_rec(T + h, h);
  }
  const cplx *w = table_for(L);
  for (size_t n = 0; n < h; n++) {
    cplx a = T[n], b = T[n + h];
    double cr = b.re * w[n].re + b.im * w[n].im;
    double ci = b.im * w[n].re - b.re * w[n].im;
    T[n].re = a.re + cr;
    T[n].im = a.im + ci;
    T[n + h].re = a.re - cr;
    T[n + h].im = a.im - ci;
  }
}

/* Inside a parallel region (a task of the e pipeline's trees) the transform's
 * own tasks go to that team, as Cilk's spawns go to its workers; outside one
 * a team is started. */
static int in_team(void) {
#ifdef _OPENMP
  return omp_in_parallel();
#else
  return 1;
#endif
}

static void run_top(void (*fn)(cplx *, size_t), cplx *T, size_t L) {
  if (L < SPAWN_LENGTH || in_team()) {
    fn(T, L);
    return;
  }
#pragma omp parallel
#pragma omp single
  fn(T, L);
}

/* fft/fft.c:395-414 without the Bailey debug block. */
void ref_fft_forward(double *z, size_t length) {
  if (length < 2) return;
  ref_ensure_table(length);
  run_top(fwd_rec, (cplx *)z, length);
}

/* fft/fft.c:417-423 */
void ref_fft_inverse(double *z, size_t length) {
  if (length < 2) return;
  ref_ensure_table(length);
  run_top(inv_rec, (cplx *)z, length);
}

/* fft/pointwise.c:9-20: a_i <- a_i * b_i. */
void ref_fft_pointwise(double *a, const double *b, size_t length) {
  cplx *A = (cplx *)a;
  const cplx *B = (const cplx *)b;
  if (length >= SPAWN_LENGTH && in_team()) {
#pragma omp taskloop num_tasks(64)
    for (size_t i = 0; i < length; i++) {
      double re = A[i].re * B[i].re - A[i].im * B[i].im;
      double im = A[i].im * B[i].re + A[i].re * B[i].im;
      A[i].re = re;
      A[i].im = im;
    }
    return;
  }
#pragma omp parallel for schedule(static) if (length >= SPAWN_LENGTH)
  for (size_t i = 0; i < length; i++) {
    double re = A[i].re * B[i].re - A[i].im * B[i].im;
    double im = A[i].im * B[i].re + A[i].re * B[i].im;
    A[i].re = re;
    A[i].im = im;
  }
}

/* ------------------------------------------------------------ conversions */

static inline uint32_t bit_mask(int bits) { return bits >= 32 ? 0xffffffffu : ((1u << bits) - 1u); }

/* fft/convert.c:45-80 with the experiments/codegen/gen.rb:94-143 loops in closed
 * form: point k = floor(A / 2^{b k}) mod 2^b for k < ceil(32 AL / b), imag 0,
 * then zeros up to `length`. Returns the number of points written. */
size_t ref_int_to_fft(double *z, size_t length, const uint32_t *A, size_t AL, int bits) {
  if (bits < 1 || bits > 30) {
    fprintf(stderr, "int_to_fft: FFT with %d bits: Not implemented\n", bits);
    abort();
  }
  size_t points = AL ? (32 * AL + (size_t)bits - 1) / (size_t)bits : 0;
  uint32_t m = bit_mask(bits);
  for (size_t k = 0; k < points; k++) {
    size_t s = (size_t)bits * k;
    size_t w = s >> 5;
    unsigned sh = (unsigned)(s & 31);
    uint64_t v = A[w];
    if (w + 1 < AL) v |= (uint64_t)A[w + 1] << 32;
    z[2 * k] = (double)((uint32_t)(v >> sh) & m);
    z[2 * k + 1] = 0.0;
  }
  for (size_t k = points; k < length; k++) {
    z[2 * k] = 0.0;
    z[2 * k + 1] = 0.0;
  }
  return points;
}

/* fft/convert.c:83-111 with the gen.rb:145-191 loops: r_k = trunc(Re(T_k)/N + 0.5),
 * words = (sum_k r_k 2^{b k}) mod 2^{32 AL}, final carry dropped. */
double ref_fft_to_int_margin(const double *z, size_t length, uint32_t *A, size_t AL, int bits) {
  if (bits < 1 || bits > 30) {
    fprintf(stderr, "fft_to_int: FFT with %d bits: Not implemented\n", bits);
    abort();
  }
  double scale = 1.0 / (double)length;
  double margin = 0.0;
  if (AL == 0) return 0.0;
  uint64_t carry = 0;
  uint32_t acc = 0;
  int pos = 0;
  size_t W = 0;
  for (size_t k = 0;; k++) {
    double f = z[2 * k] * scale;
    double d = fabs(f - rint(f));
    if (d > margin) margin = d;
    carry += (uint64_t)(int64_t)(f + 0.5);
    int need = bits;
    while (need > 0) {
      int take = need < 32 - pos ? need : 32 - pos;
      acc |= ((uint32_t)carry & bit_mask(take)) << pos;
      carry >>= take;
      pos += take;
      need -= take;
      if (pos == 32) {
        A[W++] = acc;
        acc = 0;
        pos = 0;
        if (W == AL) return margin;
      }
    }
  }
}

void ref_fft_to_int(const double *z, size_t length, uint32_t *A, size_t AL, int bits) {
  (void)ref_fft_to_int_margin(z, length, A, AL, bits);
}

/* -------------------------------------------------------------- multiplies */

/* bigfloat/mul.c:108-124: bits-per-point schedule (before the -7). */
int ref_sched_bits(size_t CL) {
  int b = 20;
  if (CL > 2000) b = 19;
  if (CL > 9000) b = 18;
  if (CL > 25000) b = 17;
  if (CL > 80000) b = 16;
  if (CL > 280000) b = 15;
  if (CL > 1000000) b = 14;
  if (CL > 4000000) b = 13;
  if (CL > 13000000) b = 12;
  if (CL > 50000000) b = 11;
  if (CL > 200000000) b = 10;
  if (CL > 800000000) b = 9;
  return b - 7;
}

/* bigfloat/mul.c:176-220: schoolbook into CL words (CT may alias AT/BT). */
void ref_basecase_mul(uint32_t *CT, size_t CL, const uint32_t *AT, size_t AL, const uint32_t *BT, size_t BL) {
  if (CT == AT || CT == BT) {
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (CL ? CL : 1));
    ref_basecase_mul(tmp, CL, AT, AL, BT, BL);
    memcpy(CT, tmp, sizeof(uint32_t) * CL);
    free(tmp);
    return;
  }
  memset(CT, 0, sizeof(uint32_t) * CL);
  if (!AL || !BL || !CL) return;
  for (size_t i = 0; i < AL; i++) {
    uint64_t carry = 0;
    size_t j = 0;
    for (; j < BL && i + j < CL; j++) {
      uint64_t v = (uint64_t)CT[i + j] + (uint64_t)AT[i] * BT[j] + carry;
      CT[i + j] = (uint32_t)v;
      carry = v >> 32;
    }
    for (size_t k = i + j; k < CL && carry; k++) {
      uint64_t v = (uint64_t)CT[k] + carry;
      CT[k] = (uint32_t)v;
      carry = v >> 32;
    }
  }
}

/* a += b (b shorter or equal), returns carry out. */
static uint32_t add_into(uint32_t *a, size_t al, const uint32_t *b, size_t bl) {
  uint64_t c = 0;
  size_t i = 0;
  for (; i < bl; i++) {
    c += (uint64_t)a[i] + b[i];
    a[i] = (uint32_t)c;
    c >>= 32;
  }
  for (; i < al && c; i++) {
    c += a[i];
    a[i] = (uint32_t)c;
    c >>= 32;
  }
  return (uint32_t)c;
}

static void sub_into(uint32_t *a, size_t al, const uint32_t *b, size_t bl) {
  int64_t br = 0;
  size_t i = 0;
  for (; i < bl; i++) {
    int64_t v = (int64_t)a[i] - (int64_t)b[i] - br;
    br = v < 0;
    a[i] = (uint32_t)v;
  }
  for (; i < al && br; i++) {
    int64_t v = (int64_t)a[i] - br;
    br = v < 0;
    a[i] = (uint32_t)v;
  }
}

/* bigfloat/mul.c:367-553: Karatsuba with the basecase below AL*BL < 50 (mul.c:13,368).
 * Pinned to the exact product (the reference's own result for valid operands). */
void ref_karatzuba_mul(uint32_t *CT, size_t CL, const uint32_t *AT, size_t AL, const uint32_t *BT, size_t BL) {
  if (AL * BL < 50 || AL < 2 || BL < 2) {
    ref_basecase_mul(CT, CL, AT, AL, BT, BL);
    return;
  }
  size_t full = AL + BL;
  uint32_t *out = (uint32_t *)calloc(full + 2, sizeof(uint32_t));
  size_t S = (AL > BL ? AL : BL) / 2;
  size_t a0 = AL < S ? AL : S, a1 = AL > S ? AL - S : 0;
  size_t b0 = BL < S ? BL : S, b1 = BL > S ? BL - S : 0;
  /* z0 = A0 B0, z2 = A1 B1 */
  uint32_t *z0 = (uint32_t *)calloc(a0 + b0 + 1, sizeof(uint32_t));
  ref_karatzuba_mul(z0, a0 + b0, AT, a0, BT, b0);
  size_t z2l = a1 + b1;
  uint32_t *z2 = (uint32_t *)calloc(z2l + 1, sizeof(uint32_t));
  if (a1 && b1) ref_karatzuba_mul(z2, z2l, AT + S, a1, BT + S, b1);
  /* (A0 + A1)(B0 + B1) */
  size_t tal = (a0 > a1 ? a0 : a1) + 1, tbl = (b0 > b1 ? b0 : b1) + 1;
  uint32_t *ta = (uint32_t *)calloc(tal, sizeof(uint32_t));
  uint32_t *tb = (uint32_t *)calloc(tbl, sizeof(uint32_t));
  memcpy(ta, AT, a0 * 4);
  ta[tal - 1] = add_into(ta, tal - 1, AT + S, a1);
  memcpy(tb, BT, b0 * 4);
  tb[tbl - 1] = add_into(tb, tbl - 1, BT + S, b1);
  uint32_t *z1 = (uint32_t *)calloc(tal + tbl, sizeof(uint32_t));
  ref_karatzuba_mul(z1, tal + tbl, ta, tal, tb, tbl);
  sub_into(z1, tal + tbl, z0, a0 + b0);
  sub_into(z1, tal + tbl, z2, z2l);
  memcpy(out, z0, (a0 + b0) * 4);
  add_into(out + 2 * S, full + 2 - 2 * S, z2, z2l);
  size_t z1l = tal + tbl;
  while (z1l && !z1[z1l - 1]) z1l--;
  add_into(out + S, full + 2 - S, z1, z1l);
  size_t n = CL < full ? CL : full;
  if (CT == AT || CT == BT) {
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (CL ? CL : 1));
    memset(tmp, 0, sizeof(uint32_t) * CL);
    memcpy(tmp, out, n * 4);
    memcpy(CT, tmp, CL * 4);
    free(tmp);
  } else {
    memset(CT, 0, sizeof(uint32_t) * CL);
    memcpy(CT, out, n * 4);
  }
  free(out); free(z0); free(z2); free(ta); free(tb); free(z1);
}

/* bigfloat/mul.c:108-174: b bits per point, N = fft_length((size_t)(32/b * CL)),
 * squaring when AT == BT && AL == BL, int_to_fft -> forward -> pointwise ->
 * inverse -> fft_to_int. Returns b. */
int ref_fft_mul_bits(uint32_t *CT, size_t CL, const uint32_t *AT, size_t AL, const uint32_t *BT, size_t BL,
                     int bits, double *margin) {
  if (bits <= 0) {
    bits = ref_sched_bits(CL);
    if (bits < 7) bits = 7;
  }
  size_t length = ref_fft_length((size_t)(32.0 / bits * (double)CL));
  size_t need = (32 * CL + (size_t)bits - 1) / (size_t)bits;
  if (need > length) length = ref_fft_length(need); /* edge case: never wrap */
  int square = (AT == BT && AL == BL);
  double *Ta = (double *)aligned_alloc(64, sizeof(cplx) * length);
  double *Tb = square ? NULL : (double *)aligned_alloc(64, sizeof(cplx) * length);
  ref_int_to_fft(Ta, length, AT, AL, bits);
  if (!square) ref_int_to_fft(Tb, length, BT, BL, bits);
  ref_ensure_table(length);
  if (!square && length >= SPAWN_LENGTH && in_team()) {
#pragma omp task
    fwd_rec((cplx *)Ta, length);
    fwd_rec((cplx *)Tb, length);
#pragma omp taskwait
  } else if (!square && length >= SPAWN_LENGTH) {
#pragma omp parallel
#pragma omp single
    {
#pragma omp task
      fwd_rec((cplx *)Ta, length);
      fwd_rec((cplx *)Tb, length);
#pragma omp taskwait
    }
  } else {
    ref_fft_forward(Ta, length);
    if (!square) ref_fft_forward(Tb, length);
  }
  ref_fft_pointwise(Ta, square ? Ta : Tb, length);
  ref_fft_inverse(Ta, length);
  double m = ref_fft_to_int_margin(Ta, length, CT, CL, bits);
  if (margin) *margin = m;
  free(Tb);
  free(Ta);
  return bits;
}

int ref_fft_mul(uint32_t *CT, size_t CL, const uint32_t *AT, size_t AL, const uint32_t *BT, size_t BL) {
  return ref_fft_mul_bits(CT, CL, AT, AL, BT, BL, 0, NULL);
}

/* bigfloat/mul.c:555-676 */
void ref_bigfloat_mul(ref_bigfloat *target, const ref_bigfloat *a, const ref_bigfloat *b, size_t p) {
  if (a->len == 0 || b->len == 0) {
    free(target->coef);
    target->coef = NULL;
    target->len = 0;
    target->exp = 0;
    target->sign = 1;
    return;
  }
  if (p == 0) p = a->len + b->len;
  else p += 2; /* BIGFLOAT_EXTRA_PRECISION, include/bigfloat.h:4 */
  int64_t ae = a->exp, be = b->exp;
  size_t AL = a->len, BL = b->len;
  const uint32_t *AT = a->coef, *BT = b->coef;
  if (AL > p) { ae += (int64_t)(AL - p); AT += AL - p; AL = p; }
  if (BL > p) { be += (int64_t)(BL - p); BT += BL - p; BL = p; }
  int sign = a->sign == b->sign;
  size_t CL = AL + BL;
  int square = (AT == BT && AL == BL);
  uint32_t *C = (uint32_t *)malloc(sizeof(uint32_t) * CL);
  if (CL == 2) {
    uint64_t r = (uint64_t)AT[0] * BT[0];
    C[0] = (uint32_t)r;
    C[1] = (uint32_t)(r >> 32);
  } else if (CL < 100) {
    ref_karatzuba_mul(C, CL, AT, AL, BT, BL);
  } else {
    ref_fft_mul(C, CL, AT, AL, square ? AT : BT, BL);
  }
  /* mantissa replacement (mul.c:608-633): realloc when aliasing and too small */
  if (target->coef == a->coef || target->coef == b->coef) {
    if (CL > target->len) target->coef = (uint32_t *)realloc(target->coef, sizeof(uint32_t) * CL);
  } else {
    free(target->coef);
    target->coef = (uint32_t *)malloc(sizeof(uint32_t) * CL);
  }
  memcpy(target->coef, C, sizeof(uint32_t) * CL);
  free(C);
  target->sign = sign;
  target->exp = ae + be;
  target->len = CL;
  /* mul.c:667-675: drop one top zero word; all of them when len > 1000 */
  if (target->coef[target->len - 1] == 0) target->len--;
  if (target->len > 1000)
    while (target->coef[target->len - 1] == 0) target->len--;
}

int ref_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void ref_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------- add, sub, rcp, div */

static void ref_zero(ref_bigfloat *t) { /* bigfloat/zero.c */
  free(t->coef);
  t->coef = NULL;
  t->len = 0;
  t->exp = 0;
  t->sign = 1;
}

/* bigfloat/addsub.c:17-28: the word of x at magnitude mag (0 outside). */
static uint32_t word_at(const ref_bigfloat *x, int64_t mag) {
  if (mag < x->exp || mag >= x->exp + (int64_t)x->len) return 0;
  return x->coef[mag - x->exp];
}

/* bigfloat/addsub.c:34-64: sign of |a| - |b|, from the top magnitudes down. */
int ref_bigfloat_ucmp(const ref_bigfloat *a, const ref_bigfloat *b) {
  const int64_t am = a->exp + (int64_t)a->len, bm = b->exp + (int64_t)b->len;
  if (am != bm) return am > bm ? 1 : -1;
  for (int64_t m = am; m >= a->exp || m >= b->exp; m--) {
    const uint32_t x = word_at(a, m), y = word_at(b, m);
    if (x != y) return x > y ? 1 : -1;
  }
  return 0;
}

/* The window [bot, top) of both operands, truncated to p (+2) words from the
 * top as in bigfloat/addsub.c:66-88 (and :119-140); *exp gets the reference's
 * exponent top - p, which is not bot when the untruncated window is shorter
 * than p (a quirk kept as it is). */
static int64_t add_window(const ref_bigfloat *a, const ref_bigfloat *b, size_t p, int64_t *bot, int64_t *exp) {
  const int64_t am = a->exp + (int64_t)a->len, bm = b->exp + (int64_t)b->len;
  const int64_t top = am > bm ? am : bm;
  int64_t lo = a->exp < b->exp ? a->exp : b->exp;
  int64_t TL = top - lo;
  if (p == 0)
    p = (size_t)TL;
  else
    p += 2; /* BIGFLOAT_EXTRA_PRECISION */
  if (TL > (int64_t)p) {
    lo = top - (int64_t)p;
    TL = (int64_t)p;
  }
  *bot = lo;
  *exp = top - (int64_t)p;
  return TL;
}

/* Replace t's mantissa by w (n words). The reference writes in place when the
 * target shares an operand's mantissa (addsub.c:90-98); here every operand
 * word is read before t changes, which equals the reference wherever its
 * in-place loop reads each word before overwriting it. */
static void set_mantissa(ref_bigfloat *t, uint32_t *w) {
  free(t->coef);
  t->coef = w;
}

/* bigfloat/addsub.c:66-117: |a| + |b| in the window, sign of a, carry out
 * appended as a top word. */
static void ref_uadd(ref_bigfloat *t, const ref_bigfloat *a, const ref_bigfloat *b, size_t p) {
  int64_t bot, exp;
  const int64_t TL = add_window(a, b, p, &bot, &exp);
  uint32_t *w = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(TL + 1));
  uint64_t carry = 0;
  for (int64_t c = 0; c < TL; c++) {
    const uint64_t v = (uint64_t)word_at(a, bot + c) + word_at(b, bot + c) + carry;
    w[c] = (uint32_t)v;
    carry = v >> 32;
  }
  size_t len = (size_t)TL;
  if (carry) w[len++] = 1;
  const int sign = a->sign;
  set_mantissa(t, w);
  t->sign = sign;
  t->exp = exp;
  t->len = len;
}

/* bigfloat/addsub.c:119-171: |a| - |b| (|a| >= |b|) in the window, sign of a,
 * zero top words stripped; an all-zero result becomes the canonical zero. */
static void ref_usub(ref_bigfloat *t, const ref_bigfloat *a, const ref_bigfloat *b, size_t p) {
  int64_t bot, exp;
  const int64_t TL = add_window(a, b, p, &bot, &exp);
  uint32_t *w = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(TL + 1));
  int64_t borrow = 0;
  for (int64_t c = 0; c < TL; c++) {
    int64_t v = (int64_t)word_at(a, bot + c) - (int64_t)word_at(b, bot + c) - borrow;
    borrow = v < 0;
    w[c] = (uint32_t)(v + (borrow ? ((int64_t)1 << 32) : 0));
  }
  size_t len = (size_t)TL;
  while (len > 0 && w[len - 1] == 0) len--;
  const int sign = a->sign;
  set_mantissa(t, w);
  t->sign = sign;
  t->exp = exp;
  t->len = len;
  if (len == 0) ref_zero(t);
}

static void ref_neg(ref_bigfloat *t) { /* bigfloat/neg.c */
  if (t->len) t->sign = !t->sign;
}

/* bigfloat/addsub.c:173-198 */
void ref_bigfloat_add(ref_bigfloat *t, const ref_bigfloat *a, const ref_bigfloat *b, size_t p) {
  if (a->sign == b->sign) {
    ref_uadd(t, a, b, p);
  } else if (ref_bigfloat_ucmp(a, b) > 0) {
    ref_usub(t, a, b, p);
  } else {
    ref_usub(t, b, a, p);
    ref_neg(t);
  }
}

/* bigfloat/addsub.c:200-225 */
void ref_bigfloat_sub(ref_bigfloat *t, const ref_bigfloat *a, const ref_bigfloat *b, size_t p) {
  if (a->sign != b->sign) {
    ref_uadd(t, a, b, p);
  } else if (ref_bigfloat_ucmp(a, b) > 0) {
    ref_usub(t, a, b, p);
  } else {
    ref_usub(t, b, a, p);
    ref_neg(t);
  }
}

/* bigfloat/rcp.c:30-70: the two-word starting value of the Newton iteration
 * from the top three words of a, in double arithmetic exactly as the
 * reference writes it (UINT32_MAX, not 2^32, as the word weight). */
static void rcp_start(ref_bigfloat *t, const ref_bigfloat *a) {
  int64_t e = a->exp;
  size_t AL = a->len;
  const uint32_t *AT = a->coef;
  if (AL > 3) {
    e += (int64_t)(AL - 3);
    AT += AL - 3;
    AL = 3;
  }
  const double U = (double)UINT32_MAX;
  double val = AT[0];
  if (AL >= 2) val += (double)AT[1] * U;
  if (AL >= 3) val += (double)AT[2] * U * U;
  val = 1. / val;
  e = -e;
  while (val < U) {
    val *= U;
    e--;
  }
  const uint64_t v = (uint64_t)val;
  uint32_t *w = (uint32_t *)malloc(2 * sizeof(uint32_t));
  w[0] = (uint32_t)(v % UINT32_MAX);
  w[1] = (uint32_t)(v / UINT32_MAX);
  const int sign = a->sign;
  set_mantissa(t, w);
  t->sign = sign;
  t->len = 2;
  t->exp = e;
}

/* bigfloat/rcp.c:15-101: Newton's iteration r1 = r0 - (r0 x - 1) r0 at
 * precision p, on the reciprocal at precision p/2 + 1 (1 -> 0, 2 -> 1). */
void ref_bigfloat_rcp(ref_bigfloat *t, const ref_bigfloat *a, size_t p) {
  if (a->len == 0) {
    fprintf(stderr, "Divide by Zero\n");
    abort();
  }
  if (p == 0) {
    rcp_start(t, a);
    return;
  }
  size_t s = p / 2 + 1;
  if (p == 1) s = 0;
  if (p == 2) s = 1;
  ref_bigfloat T = {NULL, 0, 0, 1}, one = {NULL, 0, 0, 1}, t1 = {NULL, 0, 0, 1}, t2 = {NULL, 0, 0, 1},
               t3 = {NULL, 0, 0, 1};
  ref_bigfloat_rcp(&T, a, s);
  uint32_t w1 = 1;
  one.coef = &w1;
  one.len = 1;
  ref_bigfloat_mul(&t1, a, &T, p);
  ref_bigfloat_sub(&t2, &t1, &one, p);
  ref_bigfloat_mul(&t3, &t2, &T, p);
  ref_bigfloat_sub(t, &T, &t3, p);
  free(T.coef);
  free(t1.coef);
  free(t2.coef);
  free(t3.coef);
}

/* bigfloat/div.c:14-27: a * rcp(b) at precision p (0: a->len + b->len). */
void ref_bigfloat_div(ref_bigfloat *t, const ref_bigfloat *a, const ref_bigfloat *b, size_t p) {
  if (p == 0) p = a->len + b->len;
  ref_bigfloat r = {NULL, 0, 0, 1};
  ref_bigfloat_rcp(&r, b, p);
  ref_bigfloat_mul(t, a, &r, p);
  free(r.coef);
}
