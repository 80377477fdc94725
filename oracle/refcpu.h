/*
 * refcpu -- CPU restatement of iblue/bluecrunch's FFT bigfloat multiply.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the oracle the GPU product is
 * checked against, and the CPU baseline bench.py times beside it. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product (libbluecrunch.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Parity pins (see tests/test_oracle.py):
 *   ref_fft_length, ref_table_select_raw .......... closed forms + reference table values
 *   ref_int_to_fft / ref_fft_to_int ................ test/test_fft_to_int.c, test/test_int_to_fft.c
 *                                                    vectors, test/test_conversion.c round trip
 *                                                    (glibc rand seed 17699823), closed form
 *   ref_fft_forward / ref_fft_inverse .............. test/test_fft.c vector, brute-force O(N^2)
 *                                                    DFT in long double, inverse(forward)=N*x
 *   ref_fft_pointwise .............................. closed form (complex product)
 *   ref_basecase_mul / ref_karatzuba_mul /
 *   ref_fft_mul / ref_bigfloat_mul ................. exact integer products (Python int),
 *                                                    test/test_regression.c, test/test_bigfloat_exp.c
 */
#ifndef BLUECRUNCH_REFCPU_H
#define BLUECRUNCH_REFCPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same layout as the reference bigfloat_t (include/bigfloat.h:6-13). */
typedef struct {
  uint32_t *coef;
  int64_t exp;
  size_t len;
  int sign;
} ref_bigfloat;

size_t ref_fft_length(size_t source_length);
size_t ref_table_select_raw(size_t half_length);
void ref_ensure_table(size_t length);
void ref_free_table(void);
int ref_table_size(void);

/* z is interleaved (re, im) doubles, length complex points. */
void ref_fft_forward(double *z, size_t length);
void ref_fft_inverse(double *z, size_t length);
void ref_fft_pointwise(double *a, const double *b, size_t length);

size_t ref_int_to_fft(double *z, size_t length, const uint32_t *A, size_t AL, int bits);
void ref_fft_to_int(const double *z, size_t length, uint32_t *A, size_t AL, int bits);
/* Same as ref_fft_to_int but also returns max |x - round(x)| over the points read. */
double ref_fft_to_int_margin(const double *z, size_t length, uint32_t *A, size_t AL, int bits);

int ref_sched_bits(size_t CL);
void ref_basecase_mul(uint32_t *CT, size_t CL, const uint32_t *AT, size_t AL, const uint32_t *BT, size_t BL);
void ref_karatzuba_mul(uint32_t *CT, size_t CL, const uint32_t *AT, size_t AL, const uint32_t *BT, size_t BL);
/* bits <= 0: schedule of bigfloat/mul.c:109-124 clamped below at 7 (documented deviation). */
int ref_fft_mul_bits(uint32_t *CT, size_t CL, const uint32_t *AT, size_t AL, const uint32_t *BT, size_t BL,
                     int bits, double *margin);
int ref_fft_mul(uint32_t *CT, size_t CL, const uint32_t *AT, size_t AL, const uint32_t *BT, size_t BL);
void ref_bigfloat_mul(ref_bigfloat *target, const ref_bigfloat *a, const ref_bigfloat *b, size_t p);
/* bigfloat/addsub.c, rcp.c, div.c: sums, differences, Newton reciprocal and
 * quotient with the reference's truncation rules. target's old mantissa is
 * freed (the reference leaks it or reallocs it in place). */
int ref_bigfloat_ucmp(const ref_bigfloat *a, const ref_bigfloat *b);
void ref_bigfloat_add(ref_bigfloat *target, const ref_bigfloat *a, const ref_bigfloat *b, size_t p);
void ref_bigfloat_sub(ref_bigfloat *target, const ref_bigfloat *a, const ref_bigfloat *b, size_t p);
void ref_bigfloat_rcp(ref_bigfloat *target, const ref_bigfloat *a, size_t p);
void ref_bigfloat_div(ref_bigfloat *target, const ref_bigfloat *a, const ref_bigfloat *b, size_t p);

/* refe.c: the e pipeline of main.c:38-167 and what it needs from
 * bigfloat/{mul,exp,radix,string}.c. */
void ref_bigfloat_mului(ref_bigfloat *a, uint32_t b);
void ref_bigfloat_mulu(ref_bigfloat *t, const ref_bigfloat *a, uint32_t b);
void ref_bigfloat_exp(ref_bigfloat *t, const ref_bigfloat *a, uint64_t n);
void ref_bigfloat_radix(ref_bigfloat *t, const ref_bigfloat *a);
size_t ref_bigfloat_to_string10(char *s, const ref_bigfloat *v, size_t digits);
size_t ref_e_terms(size_t p);
void ref_e_bsr(ref_bigfloat *P, ref_bigfloat *Q, uint32_t a, uint32_t b);
/* e.txt ("2." + digits decimals) into out; returns its length. ms: 4 entries
 * (binary splitting, division, radix + text, total) or NULL. */
size_t ref_e_digits(size_t digits, char *out, size_t cap, double *ms);

/* Worker threads used by the parallel regions (OpenMP). */
int ref_num_threads(void);
void ref_set_threads(int n); /* OpenMP threads of later calls (n > 0) */

#ifdef __cplusplus
}
#endif
#endif
