/*
 * nbh.h -- C ABI of the MI355X FP64 two-prime NTT multiplication engine.
 *
 * Drop-in for NativeBigInt.jl's engine pair mul_fpntt2!/sqr_fpntt2!
 * (src/fpntt.jl:982-1037), reached from mul!/sqr!/mul_bal! (src/mul.jl:134,
 * 146, 161, 389) above MUL_FPNTT_THRESHOLD = 152 / SQR_FPNTT_THRESHOLD = 160
 * limbs (src/mul.jl:24-26).  Limbs are little-endian uint64_t; r receives
 * exactly m+n limbs (2n for a square) and must not alias a or b (a may equal
 * b).  Inputs need not be top-normalised: bit lengths are taken from the top
 * limb as magnitude_bits does (src/kernels/addsub.jl:150), an upper bound.
 *
 * Every function returns NBH_OK (0) or an error code; nbh_last_error_string()
 * describes the last failure on the calling thread.  Calls are thread-safe
 * (submissions are serialised inside the library).
 */
#ifndef NBH_H
#define NBH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes.  NBH_ERR_LENGTH is the reference's ArgumentError for a
 * transform length above 315*2^34 (src/fpntt.jl:948-949). */
#define NBH_OK 0
#define NBH_ERR_LENGTH 1
#define NBH_ERR_ARG 2
#define NBH_ERR_HIP 3
#define NBH_ERR_OOM 4
#define NBH_ERR_NOT_INIT 5
#define NBH_ERR_BUDGET 6

/* Library set-up on a HIP device; idempotent.  No reference counterpart (the
 * reference's engine is a pure function with a lazily built plan cache,
 * src/fpntt.jl:273-293); the plan/twiddle cache here is per device. */
int32_t nbh_init(int32_t device);
/* Waits for all device work, then frees the plan and twiddle caches, the
 * single-product workspace and the library's streams; nbh_init may be called
 * again afterwards.  Every handle (nbh_batch, nbh_hgcd, nbh_spectrum,
 * nbh_barrett, nbh_barrett_batch, nbh_server) points into those caches, so
 * while any handle is live the call does nothing and returns NBH_ERR_ARG:
 * destroy (or release) the handles first. */
int32_t nbh_shutdown(void);
const char *nbh_last_error_string(void);

/* ntt_len (src/fpntt.jl:936-951): smallest admissible m*2^k >= T. */
int32_t nbh_ntt_len(int64_t T, int64_t *N);
/* fp_ntt_params (src/fpntt.jl:964-977): chunk width b and length N. */
int32_t nbh_ntt_params(int64_t bits_a, int64_t bits_b, int32_t *b, int64_t *N);

/* mul_fpntt2! (src/fpntt.jl:982-1014) over HOST limb buffers: r[0..m+n) =
 * a[0..m) * b[0..n).  Copies in, computes on the GPU, copies out, returns
 * when r is written. */
int32_t nbh_mul(uint64_t *r, const uint64_t *a, int64_t m, const uint64_t *b, int64_t n);
/* sqr_fpntt2! (src/fpntt.jl:1016-1037) over HOST buffers: r[0..2n) = a^2. */
int32_t nbh_sqr(uint64_t *r, const uint64_t *a, int64_t n);

/* Device-resident forms of the same: pointers are device (HBM) addresses, the
 * work is enqueued on `stream` (a hipStream_t, NULL = the library's stream)
 * and the call returns without waiting.  bits_a/bits_b are the operands'
 * magnitude_bits; pass -1 to have the library read the top limbs (a small
 * synchronous device read). */
int32_t nbh_mul_dev(uint64_t *r, const uint64_t *a, int64_t m, const uint64_t *b, int64_t n, int64_t bits_a,
                    int64_t bits_b, void *stream);
int32_t nbh_sqr_dev(uint64_t *r, const uint64_t *a, int64_t n, int64_t bits_a, void *stream);

/* nbh_mul_dev without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the same product, limbs
 * and arguments as nbh_mul_dev, computed in a static device workspace of
 * 64 MiB that the library holds on each GPU it runs on until the process
 * ends, with only plans already cached.  It calls no hipMalloc / hipFree and
 * creates nothing (profiling events included).  Returns NBH_ERR_ARG with a
 * clear message when the product needs more than 64 MiB of workspace
 * (4 spectra of N doubles plus 4 bytes per 1020 result limbs: transform
 * lengths up to about 2^21, e.g. 2^20 x 2^18 limbs) or when no plan of its
 * transform length is cached yet -- one nbh_mul_dev of the same length
 * builds it.  Every call on a GPU shares that GPU's workspace: the library
 * runs its single-product calls one after another (each waits for the end of
 * the previous one on whatever stream), so only work that overlaps outside
 * that order (e.g. a captured replay beside a call) races on it. */
int32_t nbh_mul_dev_noalloc(uint64_t *r, const uint64_t *a, int64_t m, const uint64_t *b, int64_t n,
                            int64_t bits_a, int64_t bits_b, void *stream);

/* Low short products: mullo! (src/mul.jl:305-354) r[0..k) = a[0..la) *
 * b[0..lb) mod beta^k, and sqrlo! (:360-378) r[0..k) = a[0..la)^2 mod beta^k,
 * la, lb, k >= 1.  Exactly k limbs of r are written (the reference's two
 * clobber limbs r[k..k+2) are left alone), zero-filled above a product
 * shorter than k; r must not alias a or b.  Operand limbs at or above k are
 * ignored, as the reference clamps la and lb to k.  The product of the
 * clamped operands runs in the NTT pipeline when mul!/sqr! would send it
 * there (see the batch routing below) and in the schoolbook kernel
 * otherwise; either way only the coefficients and limbs below k are
 * unpacked.  Host forms copy in and out and return when r is written; the
 * device forms are enqueued on `stream` like nbh_mul_dev, with bits_* the
 * magnitude_bits of the operands as passed (-1: read the top limb). */
int32_t nbh_mullo(uint64_t *r, const uint64_t *a, int64_t la, const uint64_t *b, int64_t lb, int64_t k);
int32_t nbh_sqrlo(uint64_t *r, const uint64_t *a, int64_t la, int64_t k);
int32_t nbh_mullo_dev(uint64_t *r, const uint64_t *a, int64_t la, const uint64_t *b, int64_t lb, int64_t k,
                      int64_t bits_a, int64_t bits_b, void *stream);
int32_t nbh_sqrlo_dev(uint64_t *r, const uint64_t *a, int64_t la, int64_t k, int64_t bits_a, void *stream);

/* Batched products (device pointers).  One job is r = a*b, or r = a^2 when
 * b == NULL.  There is no batched entry in the reference: each job is one
 * mul!/sqr! call, all of them run in a handful of launches that each cover
 * every job.  Jobs the reference sends to mul_fpntt2!/sqr_fpntt2!
 * (src/fpntt.jl:982,1016; mul!: smaller operand >= 64 limbs and m + n >= 304,
 * sqr!: n >= 160, src/mul.jl:24-26,127-164,383-389) run in the NTT
 * pipeline; the rest (its Karatsuba / mul_basecase! range) run in a
 * schoolbook kernel, one workgroup per product, beside the first NTT launch.
 * bits_* <= 0 means "read the top limb".  k > 0 makes the job a low short
 * product (mullo!/sqrlo!, as nbh_mullo_dev): r[0..k) = the product of the
 * operands clamped to k limbs, mod beta^k; k = 0 is the full product,
 * r[0..m+n). */
typedef struct {
    uint64_t *r;
    const uint64_t *a;
    const uint64_t *b;
    int64_t m, n;
    int64_t bits_a, bits_b;
    int64_t k;
} nbh_job;

typedef struct nbh_batch_s *nbh_batch;
/* Plan a batch once (plans, workspace, device tile lists). */
int32_t nbh_batch_create(const nbh_job *jobs, int64_t count, nbh_batch *out);
/* Run it (enqueue on stream); can be called repeatedly. */
int32_t nbh_batch_run(nbh_batch b, void *stream);
void nbh_batch_destroy(nbh_batch b);
/* The batch's launch sequence (every phase, with the fork/join across the
 * library's side streams) captured once into a HIP graph; replay launches the
 * graph on `stream` (capturing first if needed): the same work as
 * nbh_batch_run in one submission.  The graph holds kernels and events only. */
int32_t nbh_batch_capture(nbh_batch b);
int32_t nbh_batch_replay(nbh_batch b, void *stream);
/* Arena form for servers: the bytes a batch occupies (workspace + its job and
 * tile tables), then a batch planned inside a caller-owned device region
 * with its tables uploaded on `stream` -- no device allocation and no host
 * synchronisation, so batches can be planned while others run.  Every job
 * must carry its operands' bit lengths.  nbh_batch_release frees the host
 * side only and does not wait: the caller reuses the arena once the batch's
 * last run has completed on its stream. */
int32_t nbh_batch_arena_bytes(const nbh_job *jobs, int64_t count, int64_t *bytes);
int32_t nbh_batch_create_in(const nbh_job *jobs, int64_t count, void *arena, int64_t arena_bytes, void *stream,
                            nbh_batch *out);
void nbh_batch_release(nbh_batch b);
/* One-shot: create, run, synchronise, destroy. */
int32_t nbh_mul_batch_dev(const nbh_job *jobs, int64_t count, void *stream);
/* The same on host buffers, synchronous: every job is one mul!/sqr!/mullo!/
 * sqrlo! call of the reference (src/mul.jl:127-164,305-389, engine
 * src/fpntt.jl:982,1016) on its Memory{UInt64} limbs.  bits_* are ignored
 * (read from the top limbs).  Jobs are independent: no r may overlap another
 * job's operands or result. */
int32_t nbh_mul_batch(const nbh_job *jobs, int64_t count);

/* Serving queue (no reference counterpart: the reference multiplies one
 * product per call).  Continuous batching inside an HBM budget of its own,
 * on nstreams HIP streams of default priority.  nbh_server_submit queues a
 * job on device pointers (bits_* <= 0 reads the top limb, with a host
 * synchronisation) and returns its ticket; NBH_ERR_BUDGET if the job can
 * never fit the budget.  Each nbh_server_poll retires the micro-batches that
 * have completed, admits waiting jobs (oldest first, at most max_batch per
 * micro-batch, while their workspace fits the free arena) and launches them
 * without device allocation or host synchronisation, then reports up to cap
 * completed tickets: their results are in r.  Operands and r must stay valid
 * until the ticket is reported.  nbh_server_drain blocks until every
 * submitted job has completed (poll afterwards to collect the tickets). */
typedef struct nbh_server_s *nbh_server;
int32_t nbh_server_create(int64_t budget_bytes, int32_t nstreams, int32_t max_batch, nbh_server *out);
int32_t nbh_server_submit(nbh_server s, const nbh_job *job, int64_t *ticket);
int32_t nbh_server_poll(nbh_server s, int64_t *done, int64_t cap, int64_t *ndone);
int32_t nbh_server_drain(nbh_server s);
int32_t nbh_server_stats(nbh_server s, int64_t *pending, int64_t *in_flight, int64_t *batches);
void nbh_server_destroy(nbh_server s);

/* Device memory for callers with no GPU array type of their own (e.g. plain
 * Julia): allocate, free (after the device is idle), and synchronous copies
 * (a download waits for all device work first). */
int32_t nbh_alloc(int64_t bytes, void **ptr);
void nbh_free(void *ptr);
int32_t nbh_upload(void *dst_dev, const void *src_host, int64_t bytes);
int32_t nbh_download(void *dst_host, const void *src_dev, int64_t bytes);

/* Batched 2x2 cofactor step of HGCD (device pointers): per instance a 2x2
 * matrix of mn-limb cofactors (zero-padded) and a 2-vector of n-limb operands.
 *   NBH_HGCD_FORWARD: r0 = m00*u + m01*v, r1 = m10*u + m11*v, n+mn+1 limbs
 *                     (hgcd_mul2!, src/gcd.jl:360-367, as hgcd_matrix_mul!
 *                     and cof_mul2! use it)
 *   NBH_HGCD_ADJUST:  r0 = |m11*u - m01*v|, r1 = |m00*v - m10*u|, n+mn limbs,
 *                     with signs -1/0/+1 in s0/s1 (device int32): the M^-1
 *                     application of hgcd_matrix_adjust! (src/gcd.jl:398-437)
 *                     without the ra*2^(64p) term and for any matrix, so the
 *                     results may be negative.
 *   NBH_HGCD_ADJUST_FULL: hgcd_matrix_adjust! (src/gcd.jl:398-437) itself,
 *                     with u, v the low p = n limbs and ra, rb the nn-limb
 *                     reduced tops: r0 = ra*2^(64p) + m11*u - m01*v,
 *                     r1 = rb*2^(64p) + m00*v - m10*u, written zero-padded to
 *                     L = p + max(mn+1, nn) + 1 limbs.  The reference asserts
 *                     both are nonnegative; s0/s1 are 1 (positive), 0 (zero)
 *                     or -1 where that assertion would fail (r then holds the
 *                     magnitude).
 * The four products of every instance run as one planned product batch, the
 * adds/subtractions as one more launch.  No reference counterpart as a batch. */
#define NBH_HGCD_FORWARD 0
#define NBH_HGCD_ADJUST 1
#define NBH_HGCD_ADJUST_FULL 2
typedef struct {
    const uint64_t *m00, *m01, *m10, *m11;
    const uint64_t *u, *v;
    int64_t mn, n;
    uint64_t *r0, *r1;
    int32_t *s0, *s1;
    const uint64_t *ra, *rb;  /* NBH_HGCD_ADJUST_FULL only */
    int64_t nn;
} nbh_hgcd_job;
typedef struct nbh_hgcd_s *nbh_hgcd;
int32_t nbh_hgcd_create(const nbh_hgcd_job *jobs, int64_t count, int32_t mode, nbh_hgcd *out);
int32_t nbh_hgcd_run(nbh_hgcd h, void *stream);
void nbh_hgcd_destroy(nbh_hgcd h);

/* One level of hgcd! on the device, for many instances at once: the second
 * half of a recursion level (src/gcd.jl:546-554), which applies
 *   hgcd_matrix_adjust!(M1, a, b, p, ra, rb, nn)   (src/gcd.jl:398-437)
 *   hgcd_matrix_mul!(M, M1)                        (src/gcd.jl:372-391)
 * to buffers that stay in device memory, updated in place:
 *   a <- ra*2^(64p) + n11*a[0..p) - n01*b[0..p),  b <- rb*2^(64p) + n00*b[0..p) - n10*a[0..p),
 *       each zero-padded to L = p + max(m1n+1, nn) + 1 limbs (a and b must hold L),
 *       with sa/sb = 1 (positive), 0 (zero) or -1 where the reference's
 *       nonnegativity assertion would fail (the buffer then holds the magnitude);
 *   M <- M * M1 (hgcd_mul2! per entry), zero-padded to mn + m1n + 1 limbs
 *       (each M entry must hold that many);
 *   *n_ab = max(normlen(a), normlen(b), 1) (the adjust's return value) and
 *   *n_m = max(normlen of the M entries, 1) (the new M.n), device int64s.
 * (n00, n01, n10, n11) is M1 (m1n limbs each), (m00 .. m11) is M (mn limbs
 * each, zero-padded entries allowed).  The two updates read disjoint inputs,
 * so the level's 12 products run as one planned product batch and the six
 * sums / signed differences as one more launch.  Operands are sized by their
 * full lengths (an upper bound of their magnitude), so the buffers' contents
 * may change between runs.  The handle is an nbh_hgcd: run it with
 * nbh_hgcd_run, free it with nbh_hgcd_destroy.  No reference counterpart as
 * a batch. */
typedef struct {
    uint64_t *m00, *m01, *m10, *m11;
    int64_t mn;
    const uint64_t *n00, *n01, *n10, *n11;
    int64_t m1n;
    uint64_t *a, *b;
    int64_t p;
    const uint64_t *ra, *rb;
    int64_t nn;
    int32_t *sa, *sb;
    int64_t *n_ab, *n_m;
} nbh_hgcd_level_job;
int32_t nbh_hgcd_level_create(const nbh_hgcd_level_job *jobs, int64_t count, nbh_hgcd *out);

/* Cached forward spectra of a fixed operand.  barrett_reduce!
 * (src/barrett.jl:44-81), which powermod calls once per step
 * (src/algorithms.jl:17-107), multiplies by the same modulus m and reciprocal
 * mu every time, and each mul! re-packs and re-transforms them
 * (src/fpntt.jl:1002-1009).  A spectrum holds b's forward transforms under
 * both primes at the chunk width and length that fit every a of up to
 * bits_a_max bits (fp_ntt_params(bits_a_max, bits_b), src/fpntt.jl:964-977:
 * a shorter a keeps the exactness bound and needs no longer transform), so a
 * product against it transforms a and runs the inverse only: 4 of the 6
 * transforms of mul_fpntt2!.  The product limbs are the same as
 * nbh_mul_dev's.  No reference counterpart.
 *   nbh_spectrum_create: b (n limbs, device) is read on `stream` only; the
 *     spectrum is ready once `stream` reaches this point.  bits_b = -1 reads
 *     the top limb.  The spectrum owns 2*N doubles of device memory.
 *   nbh_mul_spec_dev: r[0..m+n) = a[0..m) * b, bits_a <= bits_a_max
 *     (else NBH_ERR_ARG), bits_a = -1 reads the top limb.  The spectrum is
 *     only read, so products against it may run concurrently.
 *   nbh_spec_batch_create: a batch (run / destroy with nbh_batch_run /
 *     nbh_batch_destroy) of products against spectra.
 *   nbh_spectrum_destroy waits for the device to go idle, then frees it. */
typedef struct nbh_spectrum_s *nbh_spectrum;
int32_t nbh_spectrum_create(const uint64_t *b, int64_t n, int64_t bits_b, int64_t bits_a_max, void *stream,
                            nbh_spectrum *out);
void nbh_spectrum_destroy(nbh_spectrum s);
int32_t nbh_spectrum_info(nbh_spectrum s, int64_t *N, int32_t *b, int64_t *n, int64_t *bits_a_max);
int32_t nbh_mul_spec_dev(uint64_t *r, const uint64_t *a, int64_t m, int64_t bits_a, nbh_spectrum s, void *stream);
typedef struct {
    uint64_t *r;
    const uint64_t *a;
    int64_t m;
    int64_t bits_a;
    nbh_spectrum s;
} nbh_spec_job;
int32_t nbh_spec_batch_create(const nbh_spec_job *jobs, int64_t count, nbh_batch *out);

/* Batched Barrett reduction (barrett_reduce!, src/barrett.jl:44-81) modulo a
 * fixed k-limb m (m[k-1] != 0) with mu = floor(beta^(2k) / m) of lmu limbs
 * (barrett_mu!, src/barrett.jl:24-31; lmu is k+1, or k+2 when m = beta^(k-1)),
 * both device buffers that must stay valid while the context lives.  The
 * context caches the spectra of mu and m; a reduction batch is planned once
 * for fixed buffers: per job r[0..k) = t[0..2k) mod m (t < beta^(2k), r must
 * not alias t).  A run is the q1*mu product batch, the q3*m product batch
 * (both against the cached spectra) and one correction launch
 * (r = (T - q3 m) mod beta^(k+1), then at most two subtractions of m).
 * Operands are taken at full length, so the buffers' contents may change
 * between runs (powermod_limbs' mulred!, src/algorithms.jl:32-55). */
typedef struct nbh_barrett_s *nbh_barrett;
typedef struct nbh_barrett_batch_s *nbh_barrett_batch;
typedef struct {
    const uint64_t *t;
    uint64_t *r;
} nbh_barrett_job;
int32_t nbh_barrett_create(const uint64_t *m, int64_t k, const uint64_t *mu, int64_t lmu, void *stream,
                           nbh_barrett *out);
void nbh_barrett_destroy(nbh_barrett b);
int32_t nbh_barrett_batch_create(nbh_barrett b, const nbh_barrett_job *jobs, int64_t count, nbh_barrett_batch *out);
int32_t nbh_barrett_batch_run(nbh_barrett_batch rb, void *stream);
void nbh_barrett_batch_destroy(nbh_barrett_batch rb);

/* Stage-level entry points over device buffers, one per reference function,
 * used by the parity tests: fp_ntt_fwd! (src/fpntt.jl:715), fp_ntt_rev!
 * (:727), fp_ntt_pack! (:744), fp_ntt_pointwise! (:770), fp_ntt_unpack2!
 * (:869, with this N's 1/N constants). */
int32_t nbh_ntt_fwd_dev(double *x, int64_t N, int32_t prime, void *stream);
int32_t nbh_ntt_rev_dev(double *x, int64_t N, int32_t prime, void *stream);
int32_t nbh_pack_dev(double *x, int64_t N, const uint64_t *limbs, int64_t n, int32_t b, int64_t nch, void *stream);
int32_t nbh_pointwise_dev(double *xa, const double *xb, int64_t N, int32_t prime, void *stream);
int32_t nbh_unpack_dev(uint64_t *r, int64_t rn, const double *x1, const double *x2, int64_t N, int64_t nconv,
                       int32_t b, void *stream);

/* Plan introspection (host copies): stage count, then per stage radix, span,
 * Q and the 8 combine constants; a twiddle row (t = 1..m-1) of Q doubles. */
int32_t nbh_plan_nstages(int64_t N, int32_t prime, int32_t *nst);
int32_t nbh_plan_stage(int64_t N, int32_t prime, int32_t i, int32_t *m, int64_t *span, int64_t *Q, double *rot8);
int32_t nbh_plan_twiddle_row(int64_t N, int32_t prime, int32_t i, int32_t t, double *out);
int32_t nbh_plan_ninv(int64_t N, int32_t prime, double *ninv, double *ninvp);

/* Schedule (how stages are grouped into HBM passes) and its tuning knobs:
 * the LDS budget of a strided group tile (doubles), the largest contiguous
 * last group, and the column range of strided tiles. */
int32_t nbh_set_tuning(int32_t grp_lds_doubles, int32_t mid_max, int32_t cmin, int32_t cmax);
int32_t nbh_get_tuning(int32_t *grp_lds_doubles, int32_t *mid_max, int32_t *cmin, int32_t *cmax);
int32_t nbh_plan_groups(int64_t N, int32_t *ngroups, int64_t *la_lb_s_c /* 4 per group */);
/* Per-length knobs (the autotuner's output): entry i, 4 knobs each in the
 * order of nbh_set_tuning, applies to lengths N <= nmax[i] (nmax increasing,
 * first match wins); longer lengths use the nbh_set_tuning default.  count 0
 * clears the table. */
int32_t nbh_set_tuning_table(const int64_t *nmax, const int32_t *knobs, int32_t count);
int32_t nbh_get_tuning_for(int64_t N, int32_t *knobs4);
/* The same for products run alone (nbh_mul, nbh_sqr, nbh_mullo, nbh_sqrlo
 * and their _dev forms): one product has to spread its transforms over the
 * whole chip, so it may want smaller tiles than a batch of many.  Lengths
 * past this table use the batch table.  The schedule file's lines that start
 * with "single" fill it at init. */
int32_t nbh_set_tuning_table_single(const int64_t *nmax, const int32_t *knobs, int32_t count);
int32_t nbh_get_tuning_for_single(int64_t N, int32_t *knobs4);

/* Per-kernel timing: when enabled, every launch of the product pipeline is
 * bracketed by HIP events on the stream it is launched on.  nbh_profile_read
 * synchronises, returns the summed milliseconds and launch counts per kernel
 * kind (0 forward strided group + pack, 1 mid, 2 inverse strided group,
 * 3 unpack, 4 carry fix, 5 cofactor-step combine, 6 schoolbook products below
 * the NTT thresholds, 7 single products run whole in one workgroup) and
 * clears them.  Instrumentation only: no reference counterpart. */
#define NBH_PROF_KINDS 8
int32_t nbh_profile_enable(int32_t on);
/* A batch step's launches (one per block shape; tiles of smaller shapes
 * join a larger one when that keeps their blocks per CU) are independent; by
 * default they fork onto 8 library streams so small shapes run beside big
 * ones.
 * nside = 0 issues them in order on the caller's stream (every launch then
 * has the GPU to itself, which is what a per-launch profile should see). */
int32_t nbh_set_streams(int32_t nside);
/* The same for one batch (-1: follow the library setting).  A serving loop
 * that already overlaps micro-batches on its own streams sets 0: each
 * fork/join costs host API calls per launch. */
int32_t nbh_batch_set_streams(nbh_batch b, int32_t nside);
int32_t nbh_profile_read(double *ms, int64_t *launches, int32_t nkinds);

/* Synchronise the library stream (or `stream`). */
int32_t nbh_synchronize(void *stream);
/* Device properties the library sized itself by. */
int32_t nbh_device_info(int32_t *cus, int64_t *lds_per_block, int64_t *hbm_bytes, char *name, int32_t name_len);

#ifdef __cplusplus
}
#endif

#endif /* NBH_H */
