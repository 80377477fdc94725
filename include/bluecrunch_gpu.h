/*
 * bluecrunch_gpu.h -- thin C ABI between the C host side (bigfloat_mul,
 * _fft_mul and the fft.h entry points) and the hand-written HIP kernels for
 * gfx950 (MI355X).
 *
 * Plain pointers and sizes only. "_dev" entry points take device pointers and
 * enqueue on the library stream (or the stream passed in); "_host" entry points
 * take host pointers and copy. Every call returns 0 on success and a negative
 * BCG_E* code on failure; bcg_last_error() describes the most recent failure.
 * All entry points are thread safe (calls are serialised per device).
 */
#ifndef BLUECRUNCH_GPU_H
#define BLUECRUNCH_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BCG_OK 0
#define BCG_E_HIP (-1)      /* a HIP runtime call failed */
#define BCG_E_ARG (-2)      /* invalid argument */
#define BCG_E_NOMEM (-3)    /* device allocation failed */
#define BCG_E_MARGIN (-4)   /* rounding margin stayed >= 0.25 at the smallest digit size */
#define BCG_E_NODEV (-5)    /* no HIP device */

/* What a multiply did: digit split, transform size, worst rounding distance. */
typedef struct {
  int bits;          /* digit size b: operands cut into b-bit digits */
  int passes;        /* HBM passes per transform (1, 2 or 3) */
  size_t points;     /* complex transform length N (right-angle: 2N digits) */
  double margin;     /* worst |x - round(x)| over all 2N output coefficients */
  int retries;       /* re-runs with a smaller digit size because margin >= 0.25 */
  int square;        /* 1 when the squaring path ran (one forward transform) */
  int group_digits;  /* carry input: digits per exact group sum (8 or 16), 0 = 64-bit coefficients */
  int group_stride;  /* words stored per group sum (0 with coefficients) */
  int chunks;        /* groups the middle passes ran in (0: whole-transform passes) */
} bcg_mul_info;

/* Device set-up: select device, create the stream, warm the kernels.
 * Called implicitly by every entry point; explicit calls are idempotent. */
int bcg_init(int device);
/* Frees device buffers, tables and the stream. */
int bcg_shutdown(void);
const char *bcg_last_error(void);
/* The library stream (hipStream_t) used when a call passes stream == NULL. */
void *bcg_stream(void);
int bcg_sync(void);

/* Device memory helpers for callers without their own allocator. */
int bcg_malloc(void **ptr, size_t bytes);
int bcg_free(void *ptr);
int bcg_memcpy_h2d(void *dst, const void *src, size_t bytes);
int bcg_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* Device-to-device copy by a plain vector copy kernel on the library stream
 * (bytes a multiple of 4); also the yardstick for HBM traffic counters. */
int bcg_memcpy_d2d(void *dst, const void *src, size_t bytes);

/* C = low CL words of A*B (little-endian base 2^32 words). A == B with
 * AL == BL selects the squaring path. C may alias A or B. Replaces the
 * reference _fft_mul (include/mul.h:3, bigfloat/mul.c:108-174). */
int bcg_mul_dev(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL,
                void *stream, bcg_mul_info *info);
int bcg_mul_host(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL,
                 bcg_mul_info *info);
/* Same, with the digit size forced (0 = automatic). For margin sweeps. */
int bcg_mul_dev_bits(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL,
                     int bits, void *stream, bcg_mul_info *info);

/* bcg_mul_dev without any allocation: no hipMalloc, hipFree or other
 * allocating call, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops). No counterpart in the reference. Its buffers
 * are one static device array of 33,694,208 bytes (32.1 MiB) per GPU, held
 * until the process ends and set up by bcg_init (which must have run), so the
 * transform is limited to 2^20 points (operands up to about 400,000 words
 * each; bcg_plan_query gives the length) and an output that aliases an
 * operand of a one-pass or schoolbook product to 16,368 words; larger
 * products fail with BCG_E_ARG. Timing events (bcg_prof_enable) are the one
 * thing it may still create. Every call on a GPU shares that array: calls that overlap on one
 * GPU (other threads, other streams) race. Same result as bcg_mul_dev. */
int bcg_mul_dev_noalloc(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL,
                        void *stream, bcg_mul_info *info);

/* Schoolbook product on the GPU (exact), used for CL < 100 like the
 * reference _karatzuba_mul/_basecase_mul branch (bigfloat/mul.c:176-553). */
int bcg_basecase_dev(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL,
                     void *stream);

/* Batched products: count independent products, word arrays concatenated.
 * a_off/b_off/c_off are word offsets into A/B/C, al/bl/cl the lengths. All
 * pointers are host arrays of descriptors; A/B/C are device words. */
int bcg_mul_batch_dev(uint32_t *C, const uint32_t *A, const uint32_t *B, const size_t *c_off, const size_t *cl,
                      const size_t *a_off, const size_t *al, const size_t *b_off, const size_t *bl, size_t count,
                      void *stream, double *worst_margin);

/* Fused multiply-add: C = low CL words of A*B + D. D may not alias C. The
 * binary-splitting step P = P0*Q1 + P1 (main.c:94-96) in one call. */
int bcg_muladd_dev(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL,
                   const uint32_t *D, size_t DL, void *stream, bcg_mul_info *info);
/* Batched multiply-add; D/d_off/dl may be NULL (then the same as
 * bcg_mul_batch_dev), dl[i] == 0 means no addend for product i. */
int bcg_muladd_batch_dev(uint32_t *C, const uint32_t *A, const uint32_t *B, const uint32_t *D, const size_t *c_off,
                         const size_t *cl, const size_t *a_off, const size_t *al, const size_t *b_off, const size_t *bl,
                         const size_t *d_off, const size_t *dl, size_t count, void *stream, double *worst_margin);

/* A batch captured once and replayed (the many small products of a
 * binary-splitting level or a Newton step, repeated): the products of one
 * bcg_muladd_batch_dev call (same arguments; pointers and lengths fixed at
 * creation) are planned, their descriptors uploaded, and every kernel launch
 * captured into one hipGraph (creation also runs the batch once).
 * bcg_batch_run replays the graph on stream (NULL: the library stream), reads
 * the rounding margins back, re-runs alone any product whose margin reached
 * 0.25, and runs the products that cannot be batched (empty or aliasing
 * operands) one call each. A, B, C, D must stay valid until
 * bcg_batch_destroy. Replays of one batch must not overlap. */
typedef struct bcg_batch_s *bcg_batch;
int bcg_batch_create(bcg_batch *batch, uint32_t *C, const uint32_t *A, const uint32_t *B, const uint32_t *D,
                     const size_t *c_off, const size_t *cl, const size_t *a_off, const size_t *al,
                     const size_t *b_off, const size_t *bl, const size_t *d_off, const size_t *dl, size_t count);
int bcg_batch_run(bcg_batch batch, void *stream, double *worst_margin);
int bcg_batch_destroy(bcg_batch batch);

/* Reference-compatible transforms on device arrays of interleaved complex
 * doubles (fft.h:15-23 semantics: forward e^{+2 pi i/N}, bit-reversed output,
 * unscaled; inverse takes bit-reversed input, natural output, unscaled). */
int bcg_fft_forward_dev(double *z, size_t length, void *stream);
int bcg_fft_inverse_dev(double *z, size_t length, void *stream);
int bcg_fft_pointwise_dev(double *a, const double *b, size_t length, void *stream);
/* int_to_fft: returns points written (ceil(32*AL/bits)) through *points. */
int bcg_int_to_fft_dev(double *z, size_t length, const uint32_t *A, size_t AL, int bits, size_t *points,
                       void *stream);
int bcg_fft_to_int_dev(const double *z, size_t length, uint32_t *A, size_t AL, int bits, double *margin,
                       void *stream);

/* Host-array variants of the above (copy in, run, copy out). */
int bcg_fft_forward_host(double *z, size_t length);
int bcg_fft_inverse_host(double *z, size_t length);
int bcg_fft_pointwise_host(double *a, const double *b, size_t length);
int bcg_int_to_fft_host(double *z, size_t length, const uint32_t *A, size_t AL, int bits, size_t *points);
int bcg_fft_to_int_host(const double *z, size_t length, uint32_t *A, size_t AL, int bits, double *margin);
int bcg_basecase_host(uint32_t *C, size_t CL, const uint32_t *A, size_t AL, const uint32_t *B, size_t BL);
/* Truncated FFT (fft/tft.c intended semantics, pinned by test/test_tft.c). */
int bcg_tft_forward_host(double *z, size_t len, int k);
int bcg_tft_inverse_host(double *z, size_t len, int k);

/* Plan inspection: the digit size and transform size a product would use. */
int bcg_plan_query(size_t AL, size_t BL, int *bits, size_t *points, int *passes);

/* Pass splits of a transform of N points (N = 2^k or 3*2^k, >= 64). Returns
 * the number of candidates; with index >= 0 also the split of that one
 * (passes, L[3] pass lengths, C[3] columns per workgroup). */
int bcg_plan_candidates(size_t N, int index, int *passes, int *L, int *C);
/* Use candidate `index` for every transform of N points (-1: back to the
 * default or the autotuned choice). */
int bcg_plan_force(size_t N, int index);
/* Plan digit sizes up to extra_bits (0..8) above the rounding bound b_max(N)
 * of every later product (0, the default: at most b_max). No counterpart in
 * the reference; it exists to exercise the rounding-margin safety net: with a
 * bias the first attempt of a product tends to reach a margin >= 0.25 and is
 * re-run with smaller digits, which bcg_stats counts as retries. */
int bcg_plan_digit_bias(int extra_bits);
/* Chunked middle passes of three-pass products: the second forward pass, the
 * middle pass and the second inverse pass run group by group of blocks, each
 * group (`bytes` of operand points) kept in the GPU's memory-side cache from
 * one of its launches to the next; 0 runs each pass over the whole transform.
 * streams (1 to 4) deals the groups round-robin over the call's stream and
 * up to three side streams (joined back before the product continues); 1
 * keeps them on one stream. Defaults: 128 MiB, 2. No counterpart in the
 * reference (a tuning knob). */
int bcg_plan_chunk(size_t bytes, int streams);
/* Index of the candidate transforms of N points use now (forced, autotuned
 * or default); negative BCG_E* code when N has no split. */
int bcg_plan_current(size_t N);
/* Autotuner: times every pass split of the transform an AL x BL product uses
 * (synthetic words, `reps` runs each) and keeps the fastest for that length.
 * times_ms (optional) receives up to max_out per-candidate milliseconds. */
int bcg_autotune(size_t AL, size_t BL, int square, int reps, int *chosen, double *times_ms, int max_out);

/* Autotuner over length classes: for every transform length class N
 * (2^k, 3*2^k) in [lo, hi], times every pass split -- and at 8192 points the
 * one-workgroup product kernel against the two-pass split -- on a batch of
 * synthetic products of that class run the way bcg_muladd_batch_dev runs them
 * (about 2^23 points per batch, at most 256 products), and keeps the fastest
 * for the batched products of that length (bcg_mul_batch_dev,
 * bcg_muladd_batch_dev, captured batches, the e pipeline's tree levels).
 * *tuned (optional) receives the number of classes with a choice.
 * bcg_plan_force_batch / bcg_plan_current_batch pin and read the choice
 * (indices into bcg_plan_candidates). */
int bcg_autotune_classes(size_t lo, size_t hi, int reps, int *tuned);
int bcg_plan_force_batch(size_t N, int index);
int bcg_plan_current_batch(size_t N);

/* e = 1 + P/Q (the reference main.c pipeline).
 * bcg_e_terms: number of series terms for `words` words of precision
 * (main.c:38-61). bcg_e_bsr_dev: P = sum_{k=1}^{T} (k+1)...T and Q = T! by
 * level-synchronous binary splitting on the GPU (main.c:63-109, same P and Q);
 * *pq receives one device allocation holding P (pl words) then Q (ql words),
 * released with bcg_free. */
size_t bcg_e_terms(size_t words);
/* F = floor(P 2^{32 nw} / Q) exactly (nw words) for P < Q (top word of Q
 * nonzero), by a Newton reciprocal of Q on the GPU and a remainder correction
 * (the e pipeline's division; bigfloat_rcp/bigfloat_div below keep the
 * reference's own truncations). *F is a device allocation released with
 * bcg_free. */
int bcg_div_fraction_dev(const uint32_t *P, size_t pl, const uint32_t *Q, size_t ql, size_t nw, uint32_t **F);
/* Base-1e9 digits (most significant first) of the binary fraction F (nf
 * device words, value F / 2^{32 nf}), `groups` of them into device `digits`
 * (the role of bigfloat_radix, bigfloat/radix.c:321-360). */
int bcg_radix_dev(const uint32_t *F, size_t nf, size_t groups, uint32_t *digits);
/* The whole e pipeline on the GPU: out = "2." + `digits` decimal digits of e,
 * truncated, NUL-terminated: the e.txt of main.c:111-167, which for a few
 * small digit counts (1-4, 10-12, 19, 20, 28) holds fewer digits, as the
 * reference's bigfloat_to_string prints them; cap >= digits + 3.
 * ms (optional, 4 entries): binary splitting, division, radix conversion,
 * total wall milliseconds. */
int bcg_e_digits(size_t digits, char *out, size_t cap, double *ms);
/* Same, with flags: BCG_E_PER_MULTIPLY runs the binary-splitting products one
 * library call per product (the plain per-multiply GPU path) instead of one
 * batched call per tree level. */
#define BCG_E_PER_MULTIPLY 1
int bcg_e_digits_ex(size_t digits, char *out, size_t cap, double *ms, int flags);
int bcg_e_bsr_dev(size_t terms, uint32_t **pq, size_t *pl, size_t *ql);

/* Bigfloats whose words live in device memory: the layout of the reference's
 * bigfloat_t (include/bigfloat.h:6-13) with coef a device pointer. value =
 * (sign ? +1 : -1) * sum coef[i] 2^{32 (i + exp)}; len == 0 is zero. The
 * operations below have the reference's semantics and give its exact words,
 * exponent and sign (truncation to p + 2 words, p = 0 full precision); the
 * target may be one of the operands, and its old mantissa is released.
 * Mantissas come from bcg_bf_set or an operation and go back with bcg_bf_free;
 * a zero-initialised bcg_bigfloat is a valid zero target. */
typedef struct {
  uint32_t *coef; /* device words */
  int64_t exp;
  size_t len;
  int sign; /* 1: positive, 0: negative */
} bcg_bigfloat;
/* x = host words (len of them), exponent and sign; x's old mantissa released. */
int bcg_bf_set(bcg_bigfloat *x, const uint32_t *words, size_t len, int64_t exp, int sign);
/* x->len words of x to host memory. */
int bcg_bf_get(const bcg_bigfloat *x, uint32_t *words);
/* Release x's mantissa; x becomes zero. */
int bcg_bf_free(bcg_bigfloat *x);
/* t = a + b, a - b (bigfloat/addsub.c:173-225), a * b (bigfloat/mul.c:555-676),
 * 1 / a (bigfloat/rcp.c:15-101), a / b (bigfloat/div.c:14-27) at precision p. */
int bcg_bf_add(bcg_bigfloat *t, const bcg_bigfloat *a, const bcg_bigfloat *b, size_t p);
int bcg_bf_sub(bcg_bigfloat *t, const bcg_bigfloat *a, const bcg_bigfloat *b, size_t p);
int bcg_bf_mul(bcg_bigfloat *t, const bcg_bigfloat *a, const bcg_bigfloat *b, size_t p);
int bcg_bf_rcp(bcg_bigfloat *t, const bcg_bigfloat *a, size_t p);
int bcg_bf_div(bcg_bigfloat *t, const bcg_bigfloat *a, const bcg_bigfloat *b, size_t p);

/* Test hook for the error paths of the e pipeline (binary splitting,
 * division, radix conversion): the nth device allocation those make from now
 * on fails as if the device were out of memory (0: off; the hook clears itself
 * when it fires). No counterpart in the reference. */
int bcg_test_fail_alloc(long nth);

/* Kernel timing with HIP events recorded on the launch stream. While enabled,
 * every kernel the engine launches is bracketed by events; bcg_prof_read
 * returns, per kernel kind, launch count and summed milliseconds. */
#define BCG_PROF_KINDS 8
#define BCG_K_FWD_PACK 0  /* first forward pass, digits read from words */
#define BCG_K_FWD 1       /* middle forward pass (3-pass plans) */
#define BCG_K_MID 2       /* last forward pass + pointwise + first inverse pass */
#define BCG_K_INV 3       /* middle inverse pass */
#define BCG_K_INV_UNPACK 4 /* last inverse pass, rounding to coefficients */
#define BCG_K_CARRY 5     /* coefficient -> word carry resolve */
#define BCG_K_OTHER 6
#define BCG_K_MIDDLE 7    /* chunked middle section of a 3-pass product: every group's second
                            forward, middle and second inverse pass, timed as one span from
                            the fork of the side streams to their join */
int bcg_prof_enable(int on);
/* Bit mask of the kinds bracketed while enabled (bit k = kind k; default all):
 * timing one kind keeps the event overhead off the others. */
int bcg_prof_select(unsigned kinds);
int bcg_prof_read(int kind, long *launches, double *ms);
int bcg_prof_reset(void);

/* Running statistics over every FFT product since the last bcg_stats_reset
 * (single, batched and those inside the e pipeline): the worst rounding
 * margin, the product count, and the digit size and margin at the longest
 * transform. No counterpart in the reference (its fft_to_int keeps no
 * margin); for reports such as the e pipeline's largest transform. */
typedef struct {
  double worst_margin;      /* max |x - round(x)| over all products */
  size_t products;          /* FFT products counted */
  size_t max_points;        /* longest transform, complex points */
  int bits_at_max;          /* digit size at that length */
  double margin_at_max;     /* worst margin at that length */
  size_t retries;           /* re-runs with a smaller digit size (margin >= 0.25) */
} bcg_stats;
int bcg_stats_reset(void);
int bcg_stats_get(bcg_stats *out);

/* Deferred margin checks, for device-resident callers that queue products
 * back to back. No counterpart in the reference (its fft_to_int checks
 * nothing). Between bcg_defer_begin() and bcg_defer_end(), the calling
 * thread's products (bcg_mul_dev, bcg_muladd_dev and the single products
 * of the e pipeline and the device bigfloats) neither read their rounding
 * margin back nor re-run at
 * >= 0.25: no host synchronisation per product, bcg_mul_info.margin reads 0.
 * Each keeps its margin on the device (sets of 64 products, read together
 * when they run out); bcg_defer_end waits for the device and returns the
 * worst margin of the scope. Results of the scope are exact only if that
 * is < 0.25; otherwise the caller redoes its work outside a scope, where
 * every product checks and re-runs itself. Products whose output aliases
 * an input, batched calls and other threads' products check at once as
 * usual. One scope at a time (BCG_E_ARG otherwise); bcg_defer_end must come
 * from the thread that opened it. */
int bcg_defer_begin(void);
int bcg_defer_end(double *worst_margin);

#ifdef __cplusplus
}
#endif
#endif
