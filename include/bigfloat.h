/*
 * bigfloat.h -- the reference bigfloat type and the entry points of it that
 * libbluecrunch provides (iblue/bluecrunch include/bigfloat.h:4-45).
 * Value = (sign ? +1 : -1) * sum_i coef[i] * 2^{32 (i + exp)}.
 */
#ifndef BLUECRUNCH_BIGFLOAT_H
#define BLUECRUNCH_BIGFLOAT_H

#include <stddef.h>
#include <stdint.h>

/* bigfloat.h:4 */
#define BIGFLOAT_EXTRA_PRECISION (2)

/* bigfloat.h:6-13 -- layout is ABI */
typedef struct {
  uint32_t *coef; /* Coefficients */
  int64_t exp;    /* Exponent */
  size_t len;     /* Length */
  int sign;       /* 1: pos, 0: neg */
} _bigfloat_struct;

typedef _bigfloat_struct bigfloat_t[1];

/* bigfloat.h:17-24 -- lifecycle (bigfloat/{zero,set,free,new,copy,alloc}.c) */
void bigfloat_zero(bigfloat_t target);
int bigfloat_iszero(const bigfloat_t target);
void bigfloat_set(bigfloat_t target, uint32_t value);
void bigfloat_free(bigfloat_t target);
void bigfloat_new(bigfloat_t target);
void bigfloat_copy(bigfloat_t target, const bigfloat_t source);
void bigfloat_alloc(bigfloat_t target, size_t size);
void bigfloat_realloc(bigfloat_t target, size_t size);

/* bigfloat.h:32-33, bigfloat/mul.c:64-102 -- a *= b in place (the top word
 * of a is cleared first: callers keep a spare word for the product), and
 * target's words = a * b over target->len words. */
void bigfloat_mului(bigfloat_t a, uint32_t b);
void bigfloat_mulu(bigfloat_t target, const bigfloat_t a, uint32_t b);

/* bigfloat.h:31, bigfloat/mul.c:555-676 -- target = a * b truncated to p words
 * (p == 0: full precision). target may alias a and/or b. */
void bigfloat_mul(bigfloat_t target, const bigfloat_t a, const bigfloat_t b, size_t p);

/* bigfloat.h:29-30,34-35 -- bigfloat/addsub.c:173-225, div.c:14-27,
 * rcp.c:15-101: sum, difference, quotient and Newton reciprocal at precision
 * p (0: full), with the reference's truncation rules and results; the words
 * are computed on the GPU (include/bluecrunch_gpu.h bcg_bf_*). */
void bigfloat_add(bigfloat_t target, const bigfloat_t a, const bigfloat_t b, size_t p);
void bigfloat_sub(bigfloat_t target, const bigfloat_t a, const bigfloat_t b, size_t p);
void bigfloat_div(bigfloat_t target, const bigfloat_t a, const bigfloat_t b, size_t p);
void bigfloat_rcp(bigfloat_t target, const bigfloat_t a, size_t p);

#endif
