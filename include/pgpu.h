/*
 * pgpu.h -- C ABI of the MI355X (gfx950) product library, libpgpu.so.
 *
 * This is the drop-in boundary for preclib's large-integer product path: the
 * host precn_t layer calls these entry points at the branches where preclib's
 * dispatcher calls its transform multipliers (src/mul_basic.cpp:259-269 and
 * 203).  All buffers are plain little-endian uint64 limb arrays; every size
 * is a limb count.  No entry point frees or reallocates caller memory.
 *
 * Status codes: 0 on success, a negative PG_E* value on failure; the text of
 * the last failure on the calling thread is returned by pg_last_error().
 * The library is internally serialized (one device context per process).
 */
#ifndef PGPU_H
#define PGPU_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
    PG_OK = 0,
    PG_EINVAL = -1,    /* bad argument (size, null pointer, capacity)        */
    PG_ENOMEM = -2,    /* device memory budget or hipMalloc exhausted        */
    PG_EDEVICE = -3,   /* HIP runtime error                                  */
    PG_ERANGE = -4,    /* operand outside the path's supported size range    */
    PG_ENOINIT = -5,   /* pg_init not called                                 */
    PG_EINTERNAL = -6  /* a checked invariant failed (result not produced)   */
};

/* Product paths.  PG_PATH_AUTO applies preclib's dispatcher structure
 * (mul_basic.cpp:238-276): schoolbook when a side is <= 24 limbs or hi <= 192,
 * blocked when hi > 2*lo, else a transform.  Where preclib's CPU thresholds
 * pick mul_fft (hi <= 6144, and every square), the GPU takes the NTT, which
 * is 2-2.5x faster there on gfx950; the product is exact either way.
 * PG_PATH_NTT / PG_PATH_FFT force the transform (mul_ntt / mul_fft
 * semantics, including mul_ntt's fall back to the FFT above 2^24 limbs or
 * 2^26 points, mul_ntt.cpp:497-506). */
enum { PG_PATH_AUTO = 0, PG_PATH_NTT = 1, PG_PATH_FFT = 2, PG_PATH_NTT_REF = 3 };
/* PG_PATH_NTT runs the transform in the library's fastest digit mode (base
 * 2^32 digits over all three of preclib's primes); PG_PATH_NTT_REF runs
 * preclib's own decomposition (base-2^16 digits, primes 469762049 and
 * 1811939329, mul_ntt.cpp:331-342, 509-513).  Both return the same product. */

/* Opens device `device` and sets the device-memory budget in bytes (0 = 90%
 * of free HBM).  Replaces nothing in preclib (it has no device); must precede
 * every other call.  Calling it again with the same device is a no-op. */
int pg_init(int device, size_t hbm_budget);
void pg_shutdown(void);
const char *pg_last_error(void);

/*
 * r = a * b for host buffers.  Replaces mul_into's transform branches
 * (src/mul_basic.cpp:259-269) and, with path NTT / FFT, the public
 * precn_t mul_ntt(a, b) (src/mul_ntt.cpp:494-518) and
 * precn_t mul_fft(a, b) (src/mul_fft.cpp:386-394).
 * A zero-length operand gives the empty product (*r_len = 0), as mul_into
 * does (mul_basic.cpp:239-242).  Otherwise the top limbs must be nonzero and
 * r_cap >= an + bn; every size runs on the GPU (sides of at most 192 limbs,
 * or one side of at most 24 under PG_PATH_AUTO, on the small-product kernel
 * that replaces mul_basic / mul_karatsuba, mul_basic.cpp:76-185).  r may
 * alias a or b.  *r_len receives the trimmed length (an+bn-1 or an+bn).  If fft_max_round is not NULL it receives the largest
 * |v - nearbyint(v)| over the FP64 coefficients of every FFT run by the call
 * (0 when none ran), the quantity COUNT_FFTS reports (mul_fft.cpp:359-382).
 */
int pg_mul(uint64_t *r, size_t r_cap, size_t *r_len,
           const uint64_t *a, size_t an, const uint64_t *b, size_t bn,
           int path, double *fft_max_round);

/* Device-resident operands.  A buffer is a limb array in HBM owned by the
 * library's pool.  There is no preclib counterpart: preclib keeps every
 * operand in malloc'd host memory (init_basic.cpp:47-135). */
typedef struct pg_buf pg_buf;
int pg_buf_alloc(pg_buf **h, size_t limbs);
int pg_buf_free(pg_buf *h);
size_t pg_buf_capacity(const pg_buf *h);
int pg_buf_upload(pg_buf *h, const uint64_t *src, size_t limbs);
int pg_buf_download(const pg_buf *h, uint64_t *dst, size_t limbs);
/* Device pointer of the buffer's limbs (for callers that manage their own
 * HIP work on the library's stream). */
uint64_t *pg_buf_ptr(pg_buf *h);

/* r = a * b with all three operands resident in HBM (same contract as
 * pg_mul; r may alias a or b; an, bn are trimmed lengths).  Asynchronous
 * unless r_len or fft_max_round is non-NULL, in which case the call waits to
 * read them.  The transform length is planned from the limb counts alone (as
 * if the top limbs were full), so no device-to-host read precedes the
 * launches. */
int pg_mul_dev(pg_buf *r, size_t *r_len, const pg_buf *a, size_t an,
               const pg_buf *b, size_t bn, int path, double *fft_max_round);

/* r = a * b on the NTT path (pg_mul_dev with PG_PATH_NTT) with no
 * allocation of any kind: no reference counterpart; it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  r, a and b are device pointers (the
 * caller's own device memory, or pg_buf_ptr of library buffers); r receives
 * an + bn limbs (the top one may be zero); an, bn are trimmed lengths.
 * Every scratch buffer comes from one static 48 MiB device array per GPU,
 * so the sizes are limited: 192 < max(an, bn) and an + bn <= 2^19 limbs
 * (transforms of at most 2^20 points); other sizes fail with PG_ERANGE.
 * Asynchronous on the library's stream (pg_sync waits); calls are ordered
 * on that stream, so they never use the array at the same time. */
int pg_mul_dev_static(uint64_t *r, const uint64_t *a, size_t an, const uint64_t *b, size_t bn);

/*
 * Many products in one submission.  preclib has no batch interface: its
 * callers multiply one pair at a time through mul_into
 * (src/mul_basic.cpp:238-276), and every job here is decomposed exactly as
 * mul_into would decompose it (small / blocked / NTT / FFT, same thresholds).
 * Products of the same transform length run in the same launch.
 */
typedef struct {
    const uint64_t *a;      /* operand limbs (trimmed)                      */
    size_t an;
    const uint64_t *b;
    size_t bn;
    uint64_t *r;            /* product, capacity r_cap >= an + bn           */
    size_t r_cap;
    size_t r_len;           /* out: trimmed product length                  */
    int path;               /* PG_PATH_*                                    */
    int status;             /* out: 0 or a PG_E* code                       */
    double fft_max_round;   /* out: largest FFT rounding distance, 0 if none */
} pg_job;

/* Host buffers in and out: uploads every operand, runs the batch, downloads
 * every product.  Returns 0 if every job succeeded. */
int pg_mul_batch(pg_job *jobs, size_t n);

/* Device-resident batches: a, b and r of every job are device pointers
 * (pg_buf_ptr).  pg_batch_plan decomposes and allocates once; pg_batch_run
 * enqueues the whole batch (captured into a hipGraph on its first run and
 * replayed afterwards); results are in place after pg_sync.  r_len is not
 * filled (the caller knows an + bn; the top limb may be zero). */
typedef struct pg_batch pg_batch;
int pg_batch_plan(pg_batch **out, const pg_job *jobs, size_t n);
int pg_batch_run(pg_batch *b);
int pg_batch_errors(pg_batch *b, double *fft_max_round, size_t n);   /* per job; waits */
int pg_batch_free(pg_batch *b);

/* Newton division, div_mulinv / mod_mulinv (src/div_newton.cpp:89-139,
 * declared at prec.hpp:151-152): q = floor(a / b), r = a mod b.  b == 0 or
 * a == 0 gives q = r = 0 (empty), as div_mulinv_impl does (div_newton.cpp:
 * 90-93).  Operands need not be trimmed.  q or r may be NULL when not wanted;
 * q needs room for an - bn + 1 limbs, r for bn.  Outputs are trimmed and the
 * lengths written to q_len / r_len.  Every product runs on the GPU; the
 * reciprocal follows newton64_reciprocal_approx (div_newton.cpp:36-68) and
 * the quotient correction div_newton.cpp:114-125.  Waits for the device. */
int pg_div(uint64_t *q, size_t q_cap, size_t *q_len, uint64_t *r, size_t r_cap, size_t *r_len,
           const uint64_t *a, size_t an, const uint64_t *b, size_t bn);
int pg_div_dev(pg_buf *q, size_t q_cap, size_t *q_len, pg_buf *r, size_t r_cap, size_t *r_len,
               const pg_buf *a, size_t an, const pg_buf *b, size_t bn);

/* precn_reciprocal_newton (src/div_newton.cpp:70-87, prec.hpp:150): the
 * exact floor(2^bits / b); 0 (empty) when b == 0 or bits < bit_length(b). */
int pg_recip(uint64_t *x, size_t x_cap, size_t *x_len, const uint64_t *b, size_t bn, size_t bits);
int pg_recip_dev(pg_buf *x, size_t x_cap, size_t *x_len, const pg_buf *b, size_t bn, size_t bits);

/* precn_sqrt (src/sqrt.cpp:71-86, prec.hpp:203): floor(sqrt(a)); x needs
 * room for an / 2 + 1 limbs.  Every division and square runs on the GPU. */
int pg_sqrt(uint64_t *x, size_t x_cap, size_t *x_len, const uint64_t *a, size_t an);

/* precn_base_convert(a, 10, ...) (src/base_convert.cpp:107-172, prec.hpp:205)
 * as text: the decimal digits of a, most significant first, no leading
 * zeros and no terminator; *len = 0 for a == 0.  out needs room for
 * an * 19.27 + 1 characters. */
int pg_to_decimal(char *out, size_t cap, size_t *len, const uint64_t *a, size_t an);

/* precn_t(std::string) (src/init_basic.cpp:66-92): the value of a decimal
 * string of n characters, most significant digit first; every character
 * outside '0'..'9' is skipped, as the reference does.  out receives *len
 * trimmed limbs (0 for zero); it needs room for n * log2(10) / 64 + 1 limbs.
 * Divide and conquer on the GPU: 19-digit chunks, then node = hi * 10^k + lo
 * level by level with batched products (the inverse of pg_to_decimal). */
int pg_from_decimal(uint64_t *out, size_t cap, size_t *len, const char *s, size_t n);

/* chud_bs(0, nterms, need_p) (pi/pi_chudnovsky.cpp:117-148): P, Q and T of
 * the Chudnovsky series over terms [0, nterms), nterms <= 2^20.  T is
 * returned as its magnitude (it is positive for every nterms >= 1).  p may
 * be NULL (need_p = false, as the root call of pi_digits, line 168). */
int pg_chud_split(uint64_t *p, size_t p_cap, size_t *p_len, uint64_t *q, size_t q_cap, size_t *q_len,
                  uint64_t *t, size_t t_cap, size_t *t_len, size_t nterms);

/* The pi program, pi_digits (pi/pi_chudnovsky.cpp:162-193): "3." followed by
 * `digits` digits (the text its --file output holds, without the final
 * newline).  phase_ms, if not NULL, receives 4 wall times in ms:
 * binary_split, sqrt_scale, final_div, to_decimal (the --phases lines). */
int pg_pi(char *out, size_t cap, size_t *len, size_t digits, double *phase_ms);

/* ---- serving: continuous batching of multiply and division requests ------
 * No counterpart in the reference (single-threaded library calls); this is
 * the request loop a serving caller puts in front of operator* and
 * div_mulinv (src/div_newton.cpp:131).  One worker thread owns the device.
 * Before every round it admits queued requests while their HBM footprint
 * fits the budget and a slot is free; a round then runs the next step of
 * every admitted request in shared launches (all products of the round in
 * one batched launch), so requests that arrive while others are on the GPU
 * join them as soon as memory and slots free up.  Results are bit-identical
 * to pg_mul / pg_div on the same operands. */
enum { PG_OP_MUL = 0, PG_OP_DIV = 1 };
typedef struct {
    int op;                      /* PG_OP_MUL: r = a * b; PG_OP_DIV: r = a / b, r2 = a % b  */
    const uint64_t *a; size_t an;   /* host operands, valid until the request finishes     */
    const uint64_t *b; size_t bn;
    uint64_t *r; size_t r_cap;   /* product (an + bn) or quotient (an - bn + 1 limbs)      */
    uint64_t *r2; size_t r2_cap; /* remainder (bn limbs) or NULL                            */
    size_t r_len, r2_len;        /* out: trimmed lengths                                    */
    int status;                  /* out: 1 while pending, then 0 or a PG_E* code           */
    double t_submit, t_admit, t_done;   /* out: seconds on pg_server_clock()               */
} pg_request;
typedef struct {
    uint64_t submitted, completed, rounds, products, max_active;
} pg_server_stats;
typedef struct pg_server pg_server;
/* hbm_budget 0 = half the library's device budget; max_active 0 = 256 slots. */
int pg_server_start(pg_server **out, size_t hbm_budget, size_t max_active);
int pg_server_submit(pg_server *s, pg_request *req);   /* does not block */
int pg_server_wait(pg_server *s, pg_request *req);     /* returns req->status */
int pg_server_drain(pg_server *s);                     /* waits for every submitted request */
int pg_server_read_stats(pg_server *s, pg_server_stats *st);
int pg_server_stop(pg_server *s);                      /* finishes queued work, joins */
double pg_server_clock(void);

/* ---- plan autotuner ----------------------------------------------------
 * No counterpart in the reference (its transforms are radix-2 sweeps with no
 * tunable plan).  Each wide-mode NTT length 2^12 .. 2^26 has NTT_VARIANTS = 16
 * plans (row / column split and column-tile width, pg_kernels.h); all give
 * the same product.  pg_ntt_autotune times every variant of 2^logn_lo ..
 * 2^logn_hi (reps runs each, products compared), keeps the fastest in the
 * library's table and reports ms[(logn - logn_lo) * 16 + variant] (-1 where
 * a variant equals the default) and chosen[logn - logn_lo].  The table is
 * read at pg_init from ntt_plans.txt beside libpgpu.so ("logn variant"
 * lines) or from the file $PGPU_NTT_PLANS names.
 * A second table holds the variants of batched launches (pg_batch_plan)
 * whose leaves hold 2^21 transform points or more in all, where throughput
 * and not one product's latency counts ("batch logn variant" lines of the
 * same file; -1: the single-product variant).  preclib_amd.tune --batch
 * chooses them on the 4096-product uneven batch. */
int pg_ntt_autotune(int logn_lo, int logn_hi, int reps, double *ms, int *chosen);
int pg_ntt_plan_get(int *variants, int n);        /* variants[logn], logn < n */
int pg_ntt_plan_set(const int *variants, int n);
int pg_ntt_batch_plan_get(int *variants, int n);  /* variants[logn] or -1, logn < n */
int pg_ntt_batch_plan_set(const int *variants, int n);

/* Waits for all queued device work. */
int pg_sync(void);

/* Diagnostic, no counterpart in the reference: bytes of page-locked host
 * memory the library holds for descriptor uploads still in flight
 * (held_bytes, at most 256 MiB plus one image between two waits) and kept
 * for reuse (spare_bytes, at most 256 MiB).  Every entry point that waits for
 * the device recycles the held images. */
int pg_host_staging(size_t *held_bytes, size_t *spare_bytes);

/* Test hook, no counterpart in the reference: perturbs the next division
 * (pg_div, pg_div_dev, pg_recip and the divisions inside pg_sqrt / pg_pi)
 * so that its corrections are exercised.  what 0: the first quotient
 * estimate q = floor(a x / 2^scale) is moved by `amount` units; what 1: the
 * Newton seed of the next reciprocal is multiplied by `amount` (an estimate
 * too large, which takes the halving branch of div_newton.cpp:57-62).  The
 * results stay exact: the quotient correction (div_newton.cpp:114-125)
 * removes any error.  One-shot; 0 clears. */
int pg_debug_perturb(int what, int64_t amount);

/* Per-stage kernel timing on the library's stream (HIP events around each
 * launch).  pg_timing_read fills up to `cap` entries and returns the number of
 * stages; names are stable identifiers such as "ntt_col_fwd". */
typedef struct {
    const char *name;
    uint64_t launches;
    double total_ms;
    double bytes;      /* algorithmic bytes moved, summed over launches */
} pg_stage_time;
int pg_timing_enable(int on);
int pg_timing_reset(void);
int pg_timing_read(pg_stage_time *out, int cap);

#ifdef __cplusplus
}
#endif
#endif
