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

/* bigfloat.h:31, bigfloat/mul.c:555-676 -- target = a * b truncated to p words
 * (p == 0: full precision). target may alias a and/or b. */
void bigfloat_mul(bigfloat_t target, const bigfloat_t a, const bigfloat_t b, size_t p);

#endif
