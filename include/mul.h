/*
 * mul.h -- the reference low-level multiply API (iblue/bluecrunch
 * include/mul.h:1-3), same signatures, run on the GPU by libbluecrunch.
 * CT receives the low CL words of AT*BT (little-endian base 2^32); CT may
 * alias AT or BT; AT == BT && AL == BL is a squaring.
 */
#ifndef BLUECRUNCH_MUL_H
#define BLUECRUNCH_MUL_H

#include <stddef.h>
#include <stdint.h>

/* mul.h:1, bigfloat/mul.c:176-220 */
void _basecase_mul(uint32_t *CT, size_t CL, uint32_t *AT, size_t AL, uint32_t *BT, size_t BL);
/* mul.h:2, bigfloat/mul.c:367-553 */
void _karatzuba_mul(uint32_t *CT, size_t CL, uint32_t *AT, size_t AL, uint32_t *BT, size_t BL);
/* mul.h:3, bigfloat/mul.c:108-174; returns the digit size (bits per point) used */
int _fft_mul(uint32_t *CT, size_t CL, uint32_t *AT, size_t AL, uint32_t *BT, size_t BL);

#endif
