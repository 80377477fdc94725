/*
 * api.c -- the reference's C API (include/fft.h, include/mul.h and
 * bigfloat_mul from include/bigfloat.h) on top of the HIP engine's C ABI
 * (include/bluecrunch_gpu.h). Host-side bookkeeping only; every transform,
 * product and carry runs on the GPU. The one exception is the pair of
 * word-by-uint32 products bigfloat_mului/bigfloat_mulu, O(len) scalar loops
 * that the reference also runs on the host and that involve no transform.
 * Errors follow the reference convention: a message on stderr, then abort().
 */
#define _GNU_SOURCE /* M_PI */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/bigfloat.h"
#include "../../include/bluecrunch_gpu.h"
#include "../../include/fft.h"
#include "../../include/mul.h"

#define MUL_THRESHOLD_FFT 100 /* bigfloat/mul.c:14 KARATZUBA_THRESH */

static void die(const char *where, int rc) {
  fprintf(stderr, "%s: GPU call failed (%d): %s\n", where, rc, bcg_last_error());
  abort();
}

#define GPU(call)                      \
  do {                                 \
    int rc_ = (call);                  \
    if (rc_ != BCG_OK) die(__func__, rc_); \
  } while (0)

/* ------------------------------------------------------------- fft tables */

complex double *twiddle_table[125];
int twiddle_table_size = 0;

/* fft/table.c:155-179 (release branch): slot of the table with half length h. */
static size_t slot_of(size_t h) {
  unsigned ones = (unsigned)__builtin_popcountl(h);
  size_t tz = (size_t)__builtin_ctzl(h);
  return ones == 1 ? 2 * (tz - 1) : 2 * tz + 1;
}

/* fft/table.c:181-187: -1 (as size_t) when the slot is beyond the built ones. */
size_t table_select(size_t length) {
  size_t idx = slot_of(length);
  if (idx > (size_t)twiddle_table_size) return (size_t)-1;
  return idx;
}

/* fft/table.c:189-209: the L/2 roots e^{+2 pi i c/L}, c < L/2, computed as
 * the reference computes them (omega = 2 pi / L in double, then libm cos and
 * sin of omega * c), into slot _table_select(L/2). A slot that already holds a
 * table keeps it (the reference builds a new one and leaks the old). */
static void build_one(size_t L) {
  size_t slot = slot_of(L / 2);
  if (slot >= 125 || twiddle_table[slot]) return;
  complex double *t = (complex double *)aligned_alloc(32, sizeof(complex double) * ((L / 2 + 1) & ~(size_t)1));
  if (!t) {
    fprintf(stderr, "fft_ensure_table: out of memory\n");
    abort();
  }
  const double omega = 2 * M_PI / (double)L;
  for (size_t c = 0; c < L / 2; c++) {
    const double angle = omega * (double)c;
    t[c] = cos(angle) + I * sin(angle);
  }
  twiddle_table[slot] = t;
}

/* fft/table.c:211-228: host tables for every 2^k and 3*2^k up to length;
 * twiddle_table_size becomes the slot count of this call (as in the
 * reference, also when it is smaller than an earlier call's). The GPU is set
 * up here too, so that later calls do not pay for device initialisation. The
 * GPU's own transforms never read these tables. */
void fft_ensure_table(size_t length) {
  size_t k = 2;
  for (;;) {
    if (2 * k > length) break;
    build_one(2 * k);
    if (3 * k > length) break;
    build_one(3 * k);
    k *= 2;
  }
  twiddle_table_size = (int)slot_of(k) + 1;
  GPU(bcg_init(0));
}

/* fft/table.c:230-234. Unlike the reference (which frees the first
 * twiddle_table_size tables and leaves the pointers and the size behind), the
 * pointers are cleared and the size reset, so table_select reports no table
 * afterwards instead of a freed one. */
void fft_free_table() {
  for (int i = 0; i < 125; i++) {
    free(twiddle_table[i]);
    twiddle_table[i] = NULL;
  }
  twiddle_table_size = 0;
}

/* ------------------------------------------------------------ transforms */

/* fft/fft.c:22-41 */
size_t fft_length(size_t source_length) {
  size_t base = 1;
  for (;;) {
    if (2 * base >= source_length) return 2 * base;
    if (3 * base >= source_length) return 3 * base;
    base *= 2;
  }
}

static int valid_len(size_t n) {
  if (n < 2) return 0;
  size_t p = n % 3 == 0 ? n / 3 : n;
  return (p & (p - 1)) == 0;
}

void fft_forward(complex double *T, size_t length) {
  if (!valid_len(length)) {
    fprintf(stderr, "fft_forward: length %zu is not 2^k or 3*2^k\n", length);
    abort();
  }
  GPU(bcg_fft_forward_host((double *)T, length));
}

void fft_inverse(complex double *T, size_t length) {
  if (!valid_len(length)) {
    fprintf(stderr, "fft_inverse: length %zu is not 2^k or 3*2^k\n", length);
    abort();
  }
  GPU(bcg_fft_inverse_host((double *)T, length));
}

/* The four-step transform is how the engine runs every transform above 8192
 * points, so the Bailey entry point is the forward transform itself. */
void baileys_forward(complex double *T, size_t length) { fft_forward(T, length); }

/* fft/fft.c:149-232: the reference's Bailey four-step, whose output order
 * equals the forward transform's (fft.h does not declare it; its test,
 * test/test_baileys.c:17, declares it itself). */
void _baileys_forward(complex double *T, size_t length) { fft_forward(T, length); }

/* fft/pointwise.c:9-20 */
void fft_pointwise(complex double *a, complex double *b, size_t len) {
  GPU(bcg_fft_pointwise_host((double *)a, (const double *)b, len));
}

/* fft/convert.c:45-80 */
size_t int_to_fft(complex double *V, size_t length, const uint32_t *A, size_t AL, int bits_per_point) {
  if (bits_per_point < 1 || bits_per_point > 30) {
    fprintf(stderr, "int_to_fft: FFT with %d bits: Not implemented\n", bits_per_point);
    abort();
  }
  size_t pts = 0;
  GPU(bcg_int_to_fft_host((double *)V, length, A, AL, bits_per_point, &pts));
  return pts;
}

/* fft/convert.c:83-111 */
void fft_to_int(const complex double *T, size_t length, uint32_t *A, size_t AL, int bits_per_point) {
  if (bits_per_point < 1 || bits_per_point > 30) {
    fprintf(stderr, "fft_to_int: FFT with %d bits: Not implemented\n", bits_per_point);
    abort();
  }
  double margin = 0;
  GPU(bcg_fft_to_int_host((const double *)T, length, A, AL, bits_per_point, &margin));
}

void tft_forward(complex double *T, size_t len, int k) { GPU(bcg_tft_forward_host((double *)T, len, k)); }

void tft_inverse(complex double *T, size_t len, int k) { GPU(bcg_tft_inverse_host((double *)T, len, k)); }

/* ------------------------------------------------------------- multiplies */

void _basecase_mul(uint32_t *CT, size_t CL, uint32_t *AT, size_t AL, uint32_t *BT, size_t BL) {
  GPU(bcg_basecase_host(CT, CL, AT, AL, BT, BL));
}

void _karatzuba_mul(uint32_t *CT, size_t CL, uint32_t *AT, size_t AL, uint32_t *BT, size_t BL) {
  if (AL + BL <= 4000)
    GPU(bcg_basecase_host(CT, CL, AT, AL, BT, BL));
  else
    GPU(bcg_mul_host(CT, CL, AT, AL, BT, BL, NULL));
}

int _fft_mul(uint32_t *CT, size_t CL, uint32_t *AT, size_t AL, uint32_t *BT, size_t BL) {
  bcg_mul_info info;
  GPU(bcg_mul_host(CT, CL, AT, AL, BT, BL, &info));
  return info.bits;
}

/* ------------------------------------------------------------- lifecycle */

/* bigfloat/zero.c */
void bigfloat_zero(bigfloat_t target) {
  target->exp = 0;
  target->len = 0;
  target->sign = 1;
  free(target->coef);
  target->coef = NULL;
}

int bigfloat_iszero(const bigfloat_t target) { return target->len == 0; }

/* bigfloat/free.c */
void bigfloat_free(bigfloat_t target) {
  free(target->coef);
  target->coef = NULL;
}

/* bigfloat/new.c */
void bigfloat_new(bigfloat_t target) {
  target->sign = 1;
  target->exp = 0;
  target->len = 0;
  target->coef = NULL;
}

/* bigfloat/set.c */
void bigfloat_set(bigfloat_t target, uint32_t value) {
  bigfloat_free(target);
  if (value == 0) {
    bigfloat_zero(target);
    return;
  }
  target->exp = 0;
  target->sign = 1;
  target->len = 1;
  target->coef = (uint32_t *)malloc(sizeof(uint32_t));
  target->coef[0] = value;
}

/* bigfloat/copy.c */
void bigfloat_copy(bigfloat_t target, const bigfloat_t source) {
  bigfloat_free(target);
  target->exp = source->exp;
  target->sign = source->sign;
  target->len = source->len;
  target->coef = (uint32_t *)malloc(sizeof(uint32_t) * (source->len ? source->len : 1));
  if (source->len) memcpy(target->coef, source->coef, sizeof(uint32_t) * source->len);
}

/* bigfloat/alloc.c */
void bigfloat_alloc(bigfloat_t target, size_t size) {
  bigfloat_free(target);
  target->exp = 0;
  target->sign = 1;
  target->len = size;
  target->coef = (uint32_t *)malloc(sizeof(uint32_t) * (size ? size : 1));
}

void bigfloat_realloc(bigfloat_t target, size_t size) {
  size_t old = target->len;
  target->len = size;
  target->coef = (uint32_t *)realloc(target->coef, sizeof(uint32_t) * (size ? size : 1));
  if (size > old) memset(target->coef + old, 0, sizeof(uint32_t) * (size - old));
}

/* ------------------------------------------- products by one 32-bit word */

/* bigfloat/mul.c:64-81: a *= b in place over a->len words. The caller keeps a
 * spare top word for the product (the reference clears it first, then the
 * final carry out of the top word is dropped). O(len) scalar work on host
 * words, the same as the reference's: no transform is involved. */
void bigfloat_mului(bigfloat_t a, uint32_t b) {
  a->coef[a->len - 1] = 0;
  uint64_t carry = 0;
  for (size_t i = 0; i < a->len; i++) {
    const uint64_t v = (uint64_t)a->coef[i] * b + carry;
    a->coef[i] = (uint32_t)v;
    carry = v >> 32;
  }
}

/* bigfloat/mul.c:83-102: target's first target->len words = a * b (words of a
 * past its length read as zero; the carry out of the last word is dropped).
 * target's length, exponent and sign are left as they are. */
void bigfloat_mulu(bigfloat_t target, const bigfloat_t a, uint32_t b) {
  uint64_t carry = 0;
  for (size_t i = 0; i < target->len; i++) {
    uint64_t v = carry;
    if (i < a->len) v += (uint64_t)a->coef[i] * b;
    target->coef[i] = (uint32_t)v;
    carry = v >> 32;
  }
}

/* ---------------------------------------------------------- bigfloat_mul */

/* bigfloat/mul.c:555-676: zero check, precision p (default AL+BL, else
 * p + BIGFLOAT_EXTRA_PRECISION), low-word truncation into the exponent,
 * sign and exponent, mantissa (re)allocation, the product, then top-zero
 * normalisation. The product runs on the GPU for every size: the schoolbook
 * kernel below KARATZUBA_THRESH words, the FFT engine above. */
void bigfloat_mul(bigfloat_t target, const bigfloat_t a, const bigfloat_t b, size_t p) {
  if (bigfloat_iszero(a) || bigfloat_iszero(b)) {
    bigfloat_zero(target);
    return;
  }
  if (p == 0)
    p = a->len + b->len;
  else
    p += BIGFLOAT_EXTRA_PRECISION;

  int64_t aexp = a->exp, bexp = b->exp;
  size_t AL = a->len, BL = b->len;
  uint32_t *AT = a->coef, *BT = b->coef;
  if (AL > p) {
    size_t chop = AL - p;
    AL = p;
    aexp += (int64_t)chop;
    AT += chop;
  }
  if (BL > p) {
    size_t chop = BL - p;
    BL = p;
    bexp += (int64_t)chop;
    BT += chop;
  }
  int sign = a->sign == b->sign;
  size_t CL = AL + BL;
  int square = (AT == BT && AL == BL);

  /* Product into a fresh buffer first: the operands may be the target. */
  uint32_t *C = (uint32_t *)malloc(sizeof(uint32_t) * CL);
  if (!C) {
    fprintf(stderr, "bigfloat_mul: out of memory\n");
    abort();
  }
  if (CL < MUL_THRESHOLD_FFT)
    GPU(bcg_basecase_host(C, CL, AT, AL, square ? AT : BT, BL));
  else
    GPU(bcg_mul_host(C, CL, AT, AL, square ? AT : BT, BL, NULL));

  if (target->coef == a->coef || target->coef == b->coef) {
    if (CL > target->len) target->coef = (uint32_t *)realloc(target->coef, sizeof(uint32_t) * CL);
  } else {
    free(target->coef);
    target->coef = (uint32_t *)malloc(sizeof(uint32_t) * CL);
  }
  memcpy(target->coef, C, sizeof(uint32_t) * CL);
  free(C);
  target->sign = sign;
  target->exp = aexp + bexp;
  target->len = CL;

  /* mul.c:667-675 */
  if (target->coef[target->len - 1] == 0) target->len--;
  if (target->len > 1000)
    while (target->coef[target->len - 1] == 0) target->len--;
}

/* ------------------------------------------- add, sub, reciprocal, quotient */

/* The operands go to the GPU as device bigfloats, the operation runs there
 * with the reference's semantics (include/bluecrunch_gpu.h bcg_bf_*), and the
 * result replaces target's mantissa. */
static void bf_up(bcg_bigfloat *d, const bigfloat_t x) {
  memset(d, 0, sizeof *d);
  d->sign = 1;
  GPU(bcg_bf_set(d, x->coef, x->len, x->exp, x->sign));
}

static void bf_down(bigfloat_t target, bcg_bigfloat *d) {
  uint32_t *w = NULL;
  if (d->len) {
    w = (uint32_t *)malloc(sizeof(uint32_t) * d->len);
    if (!w) {
      fprintf(stderr, "bigfloat: out of memory\n");
      abort();
    }
    GPU(bcg_bf_get(d, w));
  }
  free(target->coef);
  target->coef = w;
  target->len = d->len;
  target->exp = d->exp;
  target->sign = d->sign;
  GPU(bcg_bf_free(d));
}

typedef int (*bf_op2)(bcg_bigfloat *, const bcg_bigfloat *, const bcg_bigfloat *, size_t);

static void bf_binary(bf_op2 op, bigfloat_t target, const bigfloat_t a, const bigfloat_t b, size_t p) {
  bcg_bigfloat A, B, T;
  memset(&T, 0, sizeof T);
  T.sign = 1;
  bf_up(&A, a);
  if (b != a) bf_up(&B, b);
  GPU(op(&T, &A, b != a ? &B : &A, p));
  GPU(bcg_bf_free(&A));
  if (b != a) GPU(bcg_bf_free(&B));
  bf_down(target, &T);
}

/* bigfloat/addsub.c:173-198 */
void bigfloat_add(bigfloat_t target, const bigfloat_t a, const bigfloat_t b, size_t p) {
  bf_binary(bcg_bf_add, target, a, b, p);
}

/* bigfloat/addsub.c:200-225 */
void bigfloat_sub(bigfloat_t target, const bigfloat_t a, const bigfloat_t b, size_t p) {
  bf_binary(bcg_bf_sub, target, a, b, p);
}

/* bigfloat/div.c:14-27 */
void bigfloat_div(bigfloat_t target, const bigfloat_t a, const bigfloat_t b, size_t p) {
  bf_binary(bcg_bf_div, target, a, b, p);
}

/* bigfloat/rcp.c:15-101 (a zero operand: "Divide by Zero" and abort, as the
 * reference) */
void bigfloat_rcp(bigfloat_t target, const bigfloat_t a, size_t p) {
  if (a->len == 0) {
    fprintf(stderr, "Divide by Zero\n");
    abort();
  }
  bcg_bigfloat A, T;
  memset(&T, 0, sizeof T);
  T.sign = 1;
  bf_up(&A, a);
  GPU(bcg_bf_rcp(&T, &A, p));
  GPU(bcg_bf_free(&A));
  bf_down(target, &T);
}
