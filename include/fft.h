/*
 * fft.h -- the reference FFT API (iblue/bluecrunch include/fft.h:5-35), same
 * names and C signatures, implemented by libbluecrunch on the GPU.
 *
 * Data-order contract (kept from the reference):
 *   fft_forward: in place, X_k = sum_j x_j e^{+2 pi i jk/N}, output in
 *                bit-reversed order (radix-3 digit innermost for 3*2^k),
 *                unscaled                              (fft/fft.c:395-414)
 *   fft_inverse: bit-reversed input, conjugate twiddles, natural output,
 *                unscaled (N * x)                      (fft/fft.c:417-423)
 *   int_to_fft:  b-bit little-endian slices of the words as real parts,
 *                zeros up to length; returns points written (fft/convert.c:45-80)
 *   fft_to_int:  round Re(T_k)/length, carry-pack into AL words
 *                                                      (fft/convert.c:83-111)
 * Arrays are host memory; each call copies to the GPU and back.
 * Unrecoverable errors print to stderr and abort(), like the reference.
 */
#ifndef BLUECRUNCH_FFT_H
#define BLUECRUNCH_FFT_H

#include <complex.h>
#include <stddef.h>
#include <stdint.h>

/* fft.h:5-6 -- host copies of e^{+2 pi i c/L}, c < L/2, per table slot. */
extern complex double *twiddle_table[125];
extern int twiddle_table_size;

/* fft.h:9-11 */
void fft_ensure_table(size_t);
size_t table_select(size_t length);
void fft_free_table();

/* fft.h:14-16 */
size_t fft_length(size_t source_length);
void fft_forward(complex double *, size_t);
void fft_inverse(complex double *, size_t);

/* fft.h:19 */
void fft_pointwise(complex double *, complex double *, size_t);

/* fft.h:22-23 */
size_t int_to_fft(complex double *, size_t, const uint32_t *, size_t, int);
void fft_to_int(const complex double *, size_t, uint32_t *, size_t, int);

/* fft.h:26-27 -- truncated FFT (fft/tft.c, semantics of test/test_tft.c) */
void tft_forward(complex double *, size_t, int);
void tft_inverse(complex double *, size_t, int);

/* fft.h:31 -- four-step transform; same result and order as fft_forward */
void baileys_forward(complex double *, size_t);

/* fft/fft.c:149-232 -- not in the reference fft.h (test/test_baileys.c:17
 * declares it); the same transform and order as fft_forward. */
void _baileys_forward(complex double *, size_t);

#endif
