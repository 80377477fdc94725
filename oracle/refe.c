/*
 * refe -- CPU restatement of the reference's e pipeline (main.c:38-167):
 * binary splitting of sum 1/k! (main.c:63-109), the quotient and the final
 * sum (bigfloat/div.c, addsub.c through refcpu.c), the conversion to base 1e9
 * (bigfloat/radix.c:1-360) and the decimal text (bigfloat/string.c:62-125).
 *
 * TEST INFRASTRUCTURE: the e-leg CPU baseline of bench.py and a checker for
 * tests/. Never linked into the product. Written from reading the reference;
 * no reference source is copied. Cilk Plus spawns become OpenMP tasks at the
 * reference's thresholds (1000 terms in e_BSR, 10000 digits in convert_rec).
 *
 * Documented deviation: the table of radix powers (radix.c:93-94) starts
 * empty on every conversion (the reference keeps its exponents across calls
 * in one process; one conversion per process, as in main.c, is the same).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "refcpu.h"

#define OLDBASE (4294967296)
#define NEWBASE (1000000000)
#define KT (1000)
#define CONVERT_SPLIT 10000 /* radix.c:147 CONVERT_DECOMPOSE_TRESH */
#define TABLE 128           /* radix.c:148 RADIX_CONV_TABLE_SIZE */
#define BSR_SPLIT 1000      /* main.c:84 */

/* ------------------------------------------------ bigfloat lifecycle */

static void bf_new(ref_bigfloat *x) { /* bigfloat/new.c */
  x->coef = NULL;
  x->exp = 0;
  x->len = 0;
  x->sign = 1;
}

static void bf_free(ref_bigfloat *x) { /* bigfloat/free.c */
  free(x->coef);
  x->coef = NULL;
}

static void bf_set(ref_bigfloat *x, uint32_t v) { /* bigfloat/set.c */
  bf_free(x);
  x->exp = 0;
  x->sign = 1;
  if (!v) {
    x->len = 0;
    return;
  }
  x->len = 1;
  x->coef = (uint32_t *)malloc(sizeof(uint32_t));
  x->coef[0] = v;
}

static void bf_copy(ref_bigfloat *t, const ref_bigfloat *s) { /* bigfloat/copy.c */
  bf_free(t);
  t->exp = s->exp;
  t->sign = s->sign;
  t->len = s->len;
  t->coef = (uint32_t *)malloc(sizeof(uint32_t) * (s->len ? s->len : 1));
  if (s->len) memcpy(t->coef, s->coef, sizeof(uint32_t) * s->len);
}

static void bf_alloc(ref_bigfloat *t, size_t n) { /* bigfloat/alloc.c */
  bf_free(t);
  t->exp = 0;
  t->sign = 1;
  t->len = n;
  t->coef = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
}

static void bf_realloc(ref_bigfloat *t, size_t n) { /* bigfloat/alloc.c: zero-extends */
  size_t old = t->len;
  t->len = n;
  t->coef = (uint32_t *)realloc(t->coef, sizeof(uint32_t) * (n ? n : 1));
  if (n > old) memset(t->coef + old, 0, sizeof(uint32_t) * (n - old));
}

/* bigfloat/mul.c:64-81: a *= b in place, the top word cleared first. */
void ref_bigfloat_mului(ref_bigfloat *a, uint32_t b) {
  a->coef[a->len - 1] = 0;
  uint64_t carry = 0;
  for (size_t i = 0; i < a->len; i++) {
    const uint64_t v = (uint64_t)a->coef[i] * b + carry;
    a->coef[i] = (uint32_t)v;
    carry = v >> 32;
  }
}

/* bigfloat/mul.c:83-102: t's t->len words = a * b. */
void ref_bigfloat_mulu(ref_bigfloat *t, const ref_bigfloat *a, uint32_t b) {
  uint64_t carry = 0;
  for (size_t i = 0; i < t->len; i++) {
    uint64_t v = carry;
    if (i < a->len) v += (uint64_t)a->coef[i] * b;
    t->coef[i] = (uint32_t)v;
    carry = v >> 32;
  }
}

/* bigfloat/exp.c:13-29: t = a^n by squaring (t may be a). */
void ref_bigfloat_exp(ref_bigfloat *t, const ref_bigfloat *a, uint64_t n) {
  ref_bigfloat cp;
  bf_new(&cp);
  bf_copy(&cp, a);
  bf_set(t, 1);
  while (n > 0) {
    if (n % 2 == 1) ref_bigfloat_mul(t, t, &cp, 0);
    ref_bigfloat_mul(&cp, &cp, &cp, 0);
    n /= 2;
  }
  bf_free(&cp);
}

/* ---------------------------------------------------- radix conversion */

/* radix.c:40-88: k base-1e9 digits of the fraction y0 (n words below its
 * point) by repeated multiplication, dropping the words that can no longer
 * reach the next digit. */
static void convert_trunc(ref_bigfloat *s, const ref_bigfloat *y0, size_t k, size_t n) {
  const double alpha = log(NEWBASE) / log(OLDBASE);
  ref_bigfloat t;
  bf_new(&t);
  bf_copy(&t, y0);
  bf_realloc(&t, t.len + 1);
  uint32_t *base = t.coef;
  for (size_t i = 1; i <= k; i++) {
    ref_bigfloat_mului(&t, NEWBASE);
    const size_t n0 = (size_t)((double)n - floor(alpha * (double)(i - 1)));
    const size_t n1 = (size_t)((double)n - floor(alpha * (double)i));
    s->coef[k - i] = t.coef[n0];
    t.len = n0;
    t.coef += n0 - n1;
    t.len -= n0 - n1;
    t.len += 1;
    t.coef[t.len - 1] = 0;
  }
  s->len = k;
  t.coef = base;
  bf_free(&t);
}

/* radix.c:93-94: the exponents e of the powers 1e9^e the tree needs, and the
 * powers themselves */
static size_t g_val[TABLE];
static ref_bigfloat g_pow[TABLE];

static size_t find(size_t k) { /* radix.c:96-104 */
  for (size_t i = 0; i < TABLE; i++)
    if (g_val[i] == k) return i;
  return (size_t)-1;
}

static void insert_unless_exists(size_t k) { /* radix.c:106-116 */
  for (size_t i = 0; i < TABLE; i++) {
    if (g_val[i] == k) break;
    if (g_val[i] == 0) {
      g_val[i] = k;
      break;
    }
  }
}

static int cmp_size(const void *a, const void *b) { /* radix.c:118-120 */
  return (int)(*(const size_t *)a - *(const size_t *)b);
}

/* radix.c:122-134: sorts the first (count - 1) exponents, count being the
 * index of the first empty slot (the last one inserted stays where it is). */
static void sort_table(void) {
  size_t len = 0;
  for (size_t i = 0; i < TABLE; i++)
    if (g_val[i] == 0) {
      len = i - 1;
      break;
    }
  qsort(g_val, len, sizeof(size_t), cmp_size);
}

static void build_table(size_t k) { /* radix.c:136-150 */
  if (k < KT) return;
  const size_t kh = (k + 1) / 2, kl = k - kh + 1, e = k - kl;
  build_table(kh);
  if (kl != kh) build_table(kl);
  insert_unless_exists(e);
}

/* radix.c:152-198: 1e9^e for every exponent, from e - 1 (one word product)
 * or e / 2 (a square, then one word product when e is odd). */
static void generate_factors(void) {
  bf_new(&g_pow[0]);
  bf_set(&g_pow[0], NEWBASE);
  ref_bigfloat_exp(&g_pow[0], &g_pow[0], g_val[0]);
  for (size_t i = 1; i < TABLE; i++) {
    if (g_val[i] == 0) break;
    const size_t e = g_val[i];
    size_t idx = find(e - 1);
    if (idx != (size_t)-1) {
      bf_new(&g_pow[i]);
      bf_alloc(&g_pow[i], g_pow[idx].len + 1);
      ref_bigfloat_mulu(&g_pow[i], &g_pow[idx], NEWBASE);
      continue;
    }
    idx = find(e / 2);
    if (idx != (size_t)-1) {
      bf_new(&g_pow[i]);
      ref_bigfloat_mul(&g_pow[i], &g_pow[idx], &g_pow[idx], 0);
      if (e & 1) {
        bf_realloc(&g_pow[i], g_pow[i].len + 1);
        ref_bigfloat_mului(&g_pow[i], NEWBASE);
      }
      continue;
    }
    fprintf(stderr, "did not find previous value for %zu in tbl\n", e);
    abort();
  }
}

static void free_table(void) { /* radix.c:212-220 */
  for (size_t i = 0; i < TABLE; i++) {
    bf_free(&g_pow[i]);
    if (g_val[i] == 0) break;
  }
}

/* radix.c:222-318: the scaled remainder tree. */
static void convert_rec(ref_bigfloat *s, size_t k, const ref_bigfloat *y, size_t n, size_t g) {
  if (k <= KT) {
    bf_alloc(s, k);
    convert_trunc(s, y, k, n);
    return;
  }
  const size_t kh = (k + 1) / 2, kl = k - kh + 1;
  /* n_h, n_l: words so that 4 g 1e9^k < 2^{32 n} */
  const size_t nh = (size_t)floor((double)(kh * log((double)NEWBASE) + log(4.) + log((double)g)) /
                                  log((double)OLDBASE) + 1);
  const size_t nl = (size_t)floor((double)(kl * log((double)NEWBASE) + log(4.) + log((double)g)) /
                                  log((double)OLDBASE) + 1);
  ref_bigfloat yh, yl, sh, sl;
  bf_new(&yh);
  bf_alloc(&yh, y->len - (n - nh));
  memcpy(yh.coef, y->coef + (n - nh), yh.len * sizeof(uint32_t));
  bf_new(&yl);
  ref_bigfloat_mul(&yl, y, &g_pow[find(k - kl)], 0);
  yl.len = n;
  uint32_t *yl_base = yl.coef;
  yl.coef += n - nl;
  yl.len -= n - nl;
  bf_new(&sh);
  bf_new(&sl);
  if (k > CONVERT_SPLIT) {
#pragma omp task shared(sh, yh)
    convert_rec(&sh, kh, &yh, nh, g);
    convert_rec(&sl, kl, &yl, nl, g);
#pragma omp taskwait
  } else {
    convert_rec(&sh, kh, &yh, nh, g);
    convert_rec(&sl, kl, &yl, nl, g);
  }
  /* fix-ups at the seam of the halves */
  if (sh.coef[0] == NEWBASE - 1 && sl.coef[sl.len - 1] == 0) {
    uint32_t carry = 1;
    for (size_t i = 0; i < sh.len; i++) {
      const uint64_t sum = sh.coef[i] + carry;
      carry = (uint32_t)(sum / NEWBASE);
      sh.coef[i] = (uint32_t)(sum % NEWBASE);
      if (carry == 0) break;
    }
  }
  const int sl_is_zero = sh.coef[0] == 0 && sl.coef[sl.len - 1] == NEWBASE - 1;
  /* s = floor(sh / 1e9) 1e9^kl + sl */
  bf_alloc(s, sh.len - 1 + kl);
  memset(s->coef, 0, kl * sizeof(uint32_t));
  memcpy(s->coef + kl, sh.coef + 1, (sh.len - 1) * sizeof(uint32_t));
  if (!sl_is_zero) ref_bigfloat_add(s, s, &sl, 0);
  bf_free(&sh);
  bf_free(&sl);
  bf_free(&yh);
  yl.coef = yl_base;
  bf_free(&yl);
}

/* radix.c:321-360: base 2^32 -> base 1e9 of a number with one word above
 * its point. */
void ref_bigfloat_radix(ref_bigfloat *t, const ref_bigfloat *a) {
  if (-a->exp + 1 != (int64_t)a->len) {
    fprintf(stderr, "Not implemented\n");
    abort();
  }
  const size_t n = (size_t)-a->exp;
  const size_t k = (size_t)(ceil((double)n * (log(OLDBASE) / log(NEWBASE))) - 1);
  const double gd = ceil(log((double)k) / log(OLDBASE)) + 1;
  const size_t g = (size_t)(gd > KT ? gd : KT);
  memset(g_val, 0, sizeof g_val);
  memset(g_pow, 0, sizeof g_pow);
  if (k >= KT) {
    build_table(k);
    sort_table();
    generate_factors();
  }
  bf_alloc(t, k);
#pragma omp parallel
#pragma omp single
  convert_rec(t, k, a, n, g);
  if (k >= KT) free_table();
  if (t->coef[t->len - 1] > NEWBASE) {
    bf_realloc(t, t->len + 1);
    t->coef[t->len - 1] = t->coef[t->len - 2] / NEWBASE;
    t->coef[t->len - 2] = t->coef[t->len - 2] % NEWBASE;
  }
  t->exp = -(int64_t)t->len + 1;
}

/* ------------------------------------------------------------ text */

static size_t int_to_str10(uint32_t v, char *s) { /* string.c:26-34 */
  for (int i = 8; i >= 0; i--) {
    s[i] = (char)('0' + v % 10);
    v /= 10;
  }
  return 9;
}

/* string.c:62-125, base 10 with an integer part of 2 (the only case the
 * reference implements, string.c:50-59). */
size_t ref_bigfloat_to_string10(char *string, const ref_bigfloat *v, size_t digits) {
  char *s0 = string;
  if (v->len == 0) {
    string[0] = '0';
    string[1] = '\0';
    return 1;
  }
  const int mag = (int)(v->len + v->exp);
  size_t c = v->len - 1;
  if (mag == 1) {
    if (v->coef[c] != 2) {
      fprintf(stderr, "Not implemented\n");
      abort();
    }
    *string++ = '2';
  }
  *string++ = '.';
  const size_t min_c = v->len > digits / 8 ? v->len - digits / 8 - 1 : 0;
  while (c-- > min_c) {
    string += int_to_str10(v->coef[c], string);
    if ((size_t)(string - s0) >= digits) break;
  }
  *string++ = '\0';
  if ((size_t)(string - s0) > digits) {
    s0[digits + 1] = '\0';
    return digits - 1;
  }
  return (size_t)(string - s0 - 1);
}

/* ------------------------------------------------------------- e */

/* main.c:38-61: terms for p words (sizeL is a double, compared with lgammal). */
size_t ref_e_terms(size_t p) {
  const double need = p * logl(4294967296.0L) + 1;
  size_t a = 0, b = 1;
  while (lgammal((long double)b) < need) b *= 2;
  while (b - a > 1) {
    const size_t m = (a + b) / 2;
    if (lgammal((long double)m) < need)
      a = m;
    else
      b = m;
  }
  return b + 2;
}

/* main.c:63-109: P/Q = sum_{k=a+1}^{b} a!/k!... (P0 Q1 + P1, Q0 Q1). */
void ref_e_bsr(ref_bigfloat *P, ref_bigfloat *Q, uint32_t a, uint32_t b) {
  if (b - a == 1) {
    bf_set(P, 1);
    bf_set(Q, b);
    return;
  }
  const uint32_t m = (a + b) / 2;
  ref_bigfloat P0, Q0, P1, Q1, tmp;
  bf_new(&P0);
  bf_new(&Q0);
  bf_new(&P1);
  bf_new(&Q1);
  if (b - a < BSR_SPLIT) {
    ref_e_bsr(&P0, &Q0, a, m);
    ref_e_bsr(&P1, &Q1, m, b);
  } else {
#pragma omp task shared(P0, Q0)
    ref_e_bsr(&P0, &Q0, a, m);
    ref_e_bsr(&P1, &Q1, m, b);
#pragma omp taskwait
  }
  bf_new(&tmp);
  ref_bigfloat_mul(&tmp, &P0, &Q1, 0);
  bf_free(&P0);
  ref_bigfloat_add(P, &tmp, &P1, 0);
  bf_free(&tmp);
  bf_free(&P1);
  ref_bigfloat_mul(Q, &Q0, &Q1, 0);
  bf_free(&Q1);
  bf_free(&Q0);
}

/* main.c:111-167: out = e.txt ("2." and `digits` decimals, no terminator in
 * the count); returns its length. ms (optional, 4 entries): binary
 * splitting, division and sum, radix conversion and text, total. */
size_t ref_e_digits(size_t digits, char *out, size_t cap, double *ms) {
#ifdef _OPENMP
  double t0 = omp_get_wtime();
#define NOW() omp_get_wtime()
#else
  double t0 = 0;
#define NOW() 0.0
#endif
  digits++;
  const size_t p = (digits + 8) / 9;
  const size_t terms = ref_e_terms(p);
  if ((uint32_t)terms != terms) {
    fprintf(stderr, "Limit Exceeded");
    abort();
  }
  ref_bigfloat P, Q, one, tmp, dec;
  bf_new(&P);
  bf_new(&Q);
  bf_new(&one);
  bf_new(&tmp);
  bf_new(&dec);
#pragma omp parallel
#pragma omp single
  ref_e_bsr(&P, &Q, 0, (uint32_t)terms);
  const double t1 = NOW();
  bf_set(&one, 1);
  ref_bigfloat_div(&tmp, &P, &Q, p);
  bf_free(&Q);
  ref_bigfloat_add(&P, &tmp, &one, p);
  bf_free(&tmp);
  bf_free(&one);
  const double t2 = NOW();
  const size_t output_len = digits + 2;
  ref_bigfloat_radix(&dec, &P);
  char *buf = (char *)malloc(output_len + 22);
  size_t len = ref_bigfloat_to_string10(buf, &dec, output_len);
  if (len > cap) len = cap;
  memcpy(out, buf, len);
  free(buf);
  bf_free(&dec);
  bf_free(&P);
  const double t3 = NOW();
  if (ms) {
    ms[0] = (t1 - t0) * 1e3;
    ms[1] = (t2 - t1) * 1e3;
    ms[2] = (t3 - t2) * 1e3;
    ms[3] = (t3 - t0) * 1e3;
  }
#undef NOW
  return len;
}
