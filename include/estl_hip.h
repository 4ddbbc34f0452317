/*
 * estl_hip.h -- C ABI of the gfx950 FFT hot path (libestl_hip.so).
 *
 * Plain C: pointers, sizes and status codes; no C++ or torch types cross this
 * boundary and no exception escapes it.  Every call returns ESTL_OK (0) or a negative
 * status; estl_last_error() gives the message of the calling thread's last failure.
 *
 * Conventions (those of the reference's FFTW3 path, FFT_FFTW3.h:26-28):
 *   - data is interleaved complex64 (re, im float pairs), rows of length n contiguous,
 *     row r of a batch at offset r * n (idist = odist = n);
 *   - direction ESTL_FORWARD = -1 computes X[k] = sum_t x[t] exp(-2 pi i t k / n),
 *     ESTL_BACKWARD = +1 the same with exp(+...), both unnormalised unless
 *     ESTL_FLAG_NORMALIZE asks for the 1/n factor;
 *   - n is a power of two (the reference's plan assumes it: ceilog in
 *     FFT_OpenCL_Contiguous.h:56, SimpleMath.h:7-11); other sizes are rejected.
 *
 * Deliberate divergences from FFT_OpenCL_Contiguous, the GPU backend being replaced:
 *   - the backward transform is computed for every batch row (the reference conjugates
 *     only the first row, FFT_OpenCL_Contiguous.h:100-102,134-136);
 *   - the device-memory query is 64-bit (uOpenCL.h:323-337 truncates it to int).
 */
#ifndef ESTL_HIP_H
#define ESTL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
#define ESTL_OK 0
#define ESTL_ERR_INVALID (-1)     /* bad argument: size not a power of two, null pointer, ... */
#define ESTL_ERR_HIP (-2)         /* HIP runtime error (launch, copy, allocation)           */
#define ESTL_ERR_NOMEM (-3)       /* device memory exhausted or over the admission budget   */
#define ESTL_ERR_UNSUPPORTED (-4) /* size/precision combination without a kernel            */
#define ESTL_ERR_BUSY (-5)        /* server queue full (back-pressure)                      */

/* ---------------------------------------------------------------- enums */
#define ESTL_FORWARD (-1)
#define ESTL_BACKWARD (+1)

#define ESTL_PRECISION_FP32 0      /* complex64 storage, FP32 arithmetic                       */
#define ESTL_PRECISION_FP16 1      /* half2 storage + per-row FP32 scale, FP32 arithmetic      */

#define ESTL_FLAG_NORMALIZE 0x1u   /* multiply the result by 1/n                               */
#define ESTL_FLAG_AUTOTUNE 0x2u    /* pick the plan by timing candidates on the device (cached) */
#define ESTL_FLAG_PLANAR 0x4u      /* in and out in the tile-planarized layout of              */
                                   /* PlannarizedComplexArray<5> (Complex.h:34-48,97-130):     */
                                   /* per 32 values, 32 reals then 32 imaginaries; FP32 only   */

/* convolution options (estl_conv_create) */
#define ESTL_CONV_CORRELATE 0x10u        /* multiply by conj(H): correlation, Convolution.h:125-131 */
#define ESTL_CONV_NORMALIZE_KERNEL 0x20u /* divide by the kernel sum K^[0], Convolution.h:141 */
#define ESTL_CONV_CENTER 0x40u           /* rotate the filter by (K+1)/2, Convolution.h:113-120 */
#define ESTL_CONV_FP16 0x80u             /* server: binary16 rows with per-row scales         */
#define ESTL_CONV_FULL_LENGTH 0x100u     /* transform whole rows even where overlap-save      */
                                         /* windows would serve (rows longer than 2^13 with   */
                                         /* filters of at most 4097 taps)                     */

typedef struct estl_plan_s *estl_plan_t;
typedef struct estl_ubatch_s *estl_ubatch_t;
typedef struct estl_conv_s *estl_conv_t;
typedef struct estl_server_s *estl_server_t;
typedef struct estl_plan2d_s *estl_plan2d_t;

/* Message of the calling thread's last failure ("" if none). */
const char *estl_last_error(void);

/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int estl_version(void);

/* Select the HIP device used by plans created afterwards on this thread and build its
 * resident state.  Replaces the context/queue set-up of OpenCLAlgorithm's constructor
 * (uOpenCL.h:284-321), which the reference repeats for every FFT object. */
int estl_init(int device);

/* Create a plan for `batch` rows of length n.  Replaces the constructor
 * FFT_OpenCL_Contiguous(int n, bool forward, int batch) (FFT_OpenCL_Contiguous.h:46-71,
 * plan at :55-61) and the factory call FFTFactorySpecific::newFFT (FFT.h:51).
 * direction: ESTL_FORWARD / ESTL_BACKWARD.  precision: ESTL_PRECISION_*.
 * flags: ESTL_FLAG_* bits.  Twiddle tables are built on the device (FP64, rounded to
 * FP32) and shared by all plans of the same size on that device. */
int estl_plan_create(estl_plan_t *plan, long n, long batch, int direction, int precision, unsigned flags);

/* Same as estl_plan_create with an explicit pass split: npasses global passes of radix
 * 2^log2radices[q] (each in [4, 13], summing to log2 n), each computed by workgroups
 * owning 2^log2tiles[q] adjacent columns (log2tiles may be NULL: default widths).
 * npasses = 0 picks the default plan.  The reference fixes the split to radix-16
 * passes with the remainder last (BGs, FFT_OpenCL_Contiguous.h:55-60); here it is the
 * plan space the autotuner searches. */
int estl_plan_create_split(estl_plan_t *plan, long n, long batch, int direction, int precision, unsigned flags,
                           int npasses, const int *log2radices, const int *log2tiles);

/* Autotune the plan for (n, batch, direction) on the current device: time the single-
 * launch kernel (n <= 2^13) and the pass splits near the balanced one, keep the fastest
 * (used by later estl_plan_create calls with ESTL_FLAG_AUTOTUNE; persisted to the file
 * named by ESTL_TUNE_CACHE if set).  Writes a JSON report of every candidate's time into
 * report (may be NULL) and returns its length.  Replaces the fixed radix-16 choice of
 * FFT_OpenCL_Contiguous.h:55-60. */
int estl_tune(long n, long batch, int direction, char *report, int len);

/* The same for a given precision (ESTL_PRECISION_FP16 tunes the FP16-storage plan). */
int estl_tune_precision(long n, long batch, int direction, int precision, char *report, int len);

/* Destroy a plan (FFT_OpenCL_Contiguous::~FFT_OpenCL_Contiguous, FFT_OpenCL_Contiguous.h:186-191). */
int estl_plan_destroy(estl_plan_t plan);

/* Asynchronous execute on device buffers: d_in and d_out hold batch*n interleaved
 * complex64 values (FP32 plans; FP16-storage plans run through estl_execute_fp16).
 * stream is a hipStream_t (NULL = the default stream).  d_in == d_out is allowed.
 * Replaces the pass-launch loop of FFT_OpenCL_Contiguous::execute
 * (FFT_OpenCL_Contiguous.h:108-124). */
int estl_execute_device(estl_plan_t plan, const void *d_in, void *d_out, void *stream);

/* estl_execute_device without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  A multi-pass plan's scratch comes from a
 * static 64 MiB device array per GPU instead of a per-stream buffer allocated on first
 * use, so the plan's n * batch may be at most 2^23 values (2^22 for an in-place plan of
 * an odd number of passes); larger plans are rejected with ESTL_ERR_INVALID before any
 * device work.  FP32 plans only.  All calls on one GPU share that array: calls that
 * overlap in time on one GPU (several streams or threads) race, and the caller must
 * serialise them.  Same results as estl_execute_device, bit for bit. */
int estl_execute_device_noalloc(estl_plan_t plan, const void *d_in, void *d_out, void *stream);

/* ---------------------------------------------------------------- 2-D transforms
 * `batch` arrays of ny rows of nx complex64, x fastest (FFT2D's data[y * xDim + x],
 * FFT.h:54-84, which runs one host round trip per row and per column).  Rows are one
 * batched 1-D launch; columns are Stockham passes at stride nx (no transpose; one pass
 * for ny <= 2^9, two for ny up to 2^13, so that every tile moves runs of at least 16
 * adjacent columns), so the transform makes two or three HBM round trips.
 * ESTL_FLAG_NORMALIZE scales by 1/(nx ny).  d_in == d_out is allowed. */
int estl_plan2d_create(estl_plan2d_t *plan, long nx, long ny, long batch, int direction, unsigned flags);
int estl_plan2d_execute(estl_plan2d_t plan, const void *d_in, void *d_out, void *stream);
/* Synchronous host-buffer form (copy in, transform, copy out), like FFT2D::execute. */
int estl_plan2d_execute_host(estl_plan2d_t plan, const void *h_in, void *h_out);
int estl_plan2d_launches(estl_plan2d_t plan);
int estl_plan2d_destroy(estl_plan2d_t plan);

/* ---------------------------------------------------------------- FP16 storage
 * No counterpart in the reference, which stores cl_float2 only
 * (FFT_OpenCL_Contiguous.h:47).  A buffer is batch*n binary16 (re, im) pairs plus one
 * FP32 scale per transform: true value = stored value * scale.  Arithmetic is FP32;
 * each pass of radix B stores 1/sqrt(B) of its result, so no stored value can leave
 * FP16's range for n <= 2^30. */

/* complex64 -> FP16 pairs with per-transform max-abs scaling: scale[t] = max over the
 * transform of max(|re|, |im|) (1 if all zero), stored = x / scale[t]. */
int estl_pack_fp16(const void *d_x, void *d_x16, float *d_scale, long n, long batch, void *stream);

/* FP16 pairs -> complex64: x = stored * scale[t]. */
int estl_unpack_fp16(const void *d_x16, const float *d_scale, void *d_x, long n, long batch, void *stream);

/* Execute a plan created with ESTL_PRECISION_FP16 on FP16 buffers; writes the output
 * scales (input scales * sqrt(n), times 1/n with ESTL_FLAG_NORMALIZE). */
int estl_execute_fp16(estl_plan_t plan, const void *d_in, const float *d_in_scale, void *d_out, float *d_out_scale,
                      void *stream);

/* Synchronous execute on host buffers with the reference's semantics: copy in, run,
 * copy out, wait (FFT_OpenCL_Contiguous::execute, FFT_OpenCL_Contiguous.h:97-142, and
 * FFT_FFTW3::execute, FFT_FFTW3.h:31).  The plan keeps its device buffers resident, so
 * host executes of one plan from several threads must be serialised by the caller. */
int estl_execute_host(estl_plan_t plan, const void *h_in, void *h_out);

/* Largest batch of length-n rows whose working set fits `budget_bytes` of device memory
 * (budget_bytes <= 0: the device's free memory).  64-bit replacement of
 * FFT_OpenCL_Contiguous::getMaxBatchSize / getRequiredMemory (FFT_OpenCL_Contiguous.h:43-44). */
long long estl_max_batch(long n, int precision, long long budget_bytes);

/* Bytes of device memory a plan's execution needs beyond its inputs and outputs. */
long long estl_plan_workspace_bytes(estl_plan_t plan);

/* JSON description of the plan (kind, passes, radices, kernels) into buf (NUL-terminated).
 * Returns the full length; the reference exposes the per-pass radix as getKernelInfo
 * (FFT_OpenCL_Contiguous.h:95). */
int estl_plan_describe(estl_plan_t plan, char *buf, int len);

/* Profile: run each launch of the plan on its own between two HIP events, `reps` times
 * after a warm-up (d_in/d_out NULL: the plan's own resident buffers), keep the mean per
 * launch and write a JSON report (per launch: info, ms, flops, GFLOP/s, share).  The
 * reference's kernelTimers / getTotalComputationFlops(q) / getKernelInfo(q) surface
 * (FFT_OpenCL_Contiguous.h:73-78,94-95; uOpenCL.h:281-282,339-340) as printed by
 * tests/Test.cpp:86-100.  Returns the report's length. */
int estl_plan_profile(estl_plan_t plan, const void *d_in, void *d_out, int reps, char *json, int len);

/* Flops of launch q: 5 n log2(radix_q) per transform (the rows kernel: 5 n log2 n). */
double estl_plan_kernel_flops(estl_plan_t plan, int q);

/* Mean milliseconds of launch q from the last estl_plan_profile (-1 if not profiled). */
double estl_plan_kernel_ms(estl_plan_t plan, int q);

/* Description of launch q (kernel, radix, tile). */
int estl_plan_kernel_info(estl_plan_t plan, int q, char *buf, int len);

/* Number of kernel launches one execute makes (the reference: one per pass,
 * FFT_OpenCL_Contiguous.h:108). */
int estl_plan_launches(estl_plan_t plan);

/* Fill count floats at d_dst with the seeded uniform[-1,1) stream:
 * value i = ((u >> 40) * 2^-23) - 1, u = splitmix64(splitmix64(seed) + offset + i).
 * No reference counterpart (the reference seeds rand() with time(NULL), tests/Test.cpp:189). */
int estl_fill_uniform(void *d_dst, long count, uint64_t seed, uint64_t offset, void *stream);

/* Copy `bytes` (a multiple of 16; both pointers 16-byte aligned) from d_src to d_dst
 * with 16-byte loads and stores, nontemporal if `nontemporal` is set: a plain read-once /
 * write-once kernel, the achievable-bandwidth reference the benchmark reports beside
 * the transforms.  No reference counterpart (the reference reports no bandwidth). */
int estl_stream_copy(const void *d_src, void *d_dst, long long bytes, int nontemporal, void *stream);

/* ---------------------------------------------------------------- uneven batches
 * One launch computes `count` transforms of different power-of-two sizes (2 .. 2^20),
 * transform i of sizes[i] complex64 values at element offset offsets[i] of the in and
 * out buffers (offsets NULL: packed back to back).  No reference counterpart: the
 * reference batches only equal sizes (batchCount, FFT.h:25-27) and needs one FFT object,
 * program and launch sequence per size (tests/Test.cpp:48-53).  Transforms must not
 * overlap; the buffers must hold the extent estl_ubatch_describe reports.  One object
 * holds one work-queue ticket and one scratch region: executes of the same object must
 * not overlap in time (order them on one stream); separate objects run concurrently. */
int estl_ubatch_create(estl_ubatch_t *ub, int count, const long *sizes, const long *offsets, int direction,
                       unsigned flags);

/* One launch of the work-queue kernel on `stream` (asynchronous). */
int estl_ubatch_execute(estl_ubatch_t ub, const void *d_in, void *d_out, void *stream);

/* The same launch captured once in a HIP graph per (d_in, d_out, stream) and replayed. */
int estl_ubatch_execute_graph(estl_ubatch_t ub, const void *d_in, void *d_out, void *stream);

/* JSON description (transforms, work items, workgroups, element extent) into buf. */
int estl_ubatch_describe(estl_ubatch_t ub, char *buf, int len);

int estl_ubatch_destroy(estl_ubatch_t ub);

/* ---------------------------------------------------------------- convolution
 * y[r][i] = sum_t h[t] x[r][(i - t) mod n] for `rows` real rows of length n (a power of
 * two) and a real filter of any length k (longer than n: folds modulo n), computed as
 * transform, multiply, inverse -- the pipeline of Convolution2D_FFT::convolve
 * (Convolution.h:95-148), batched over 1-D rows and fused on the device.  Two real rows
 * share one complex transform; a single FP32 row of n >= 2^14 whose input and output
 * are 8-byte aligned runs through complex transforms of n/2 values instead (its even
 * and odd samples as one complex row).  Flags: ESTL_CONV_* above. */
int estl_conv_create(estl_conv_t *conv, long n, long rows, int precision, unsigned flags);

/* Fold (and, with ESTL_CONV_CENTER, rotate) the k-value device filter d_h and keep its
 * spectrum resident on the device; later executes reuse it. */
int estl_conv_set_filter(estl_conv_t conv, const float *d_h, long k, void *stream);

/* y = x (*) h for all rows (device buffers of rows * n floats; asynchronous).  The
 * object's spectrum and scratch are its own: executes (and set_filter) of one object
 * must not overlap in time; the server (estl_server_*) gives every request its own. */
int estl_conv_execute(estl_conv_t conv, const float *d_x, float *d_y, void *stream);

/* FP16-storage convolution (object created with ESTL_PRECISION_FP16; no counterpart in
 * the reference, which stores FP32 only): d_x16 / d_y16 are rows * n binary16 reals,
 * true x = stored * d_x_scale[row] (estl_pack_fp16 with n/2 complex values per row
 * packs real rows this way); arithmetic is FP32 with 1/sqrt(radix) storage scaling per
 * pass, and the filter spectrum is divided by L = sum |h|, so every stored output is
 * in [-1, 1] and d_y_scale[row] = d_x_scale[row] * L. */
int estl_conv_execute_fp16(estl_conv_t conv, const void *d_x16, const float *d_x_scale, void *d_y16,
                           float *d_y_scale, void *stream);

/* Device bytes the object holds (scratch, spectrum). */
long long estl_conv_workspace_bytes(estl_conv_t conv);

int estl_conv_destroy(estl_conv_t conv);

/* ---------------------------------------------------------------- serving loop
 * A request loop in front of the FFT and convolution pipelines.  The reference has no
 * counterpart: its FFT objects are built and run by the caller one at a time
 * (tests/Test.cpp:53-103) and device memory is only checked by a message
 * (FFT_OpenCL_Contiguous.h:49).  Requests are admitted in arrival order by HBM
 * footprint against `hbm_budget_bytes` (one arena allocated at create), launched on a
 * pool of `streams` default-priority streams while earlier ones still run, and their
 * workspace returns to the arena as soon as they complete.  A request whose footprint
 * exceeds the whole budget fails with ESTL_ERR_NOMEM; a full queue returns
 * ESTL_ERR_BUSY.  The server's streams wait for the work queued on `input_stream` at
 * submit (the producer of the inputs); input buffers must stay valid until the request
 * is DONE.  Request groups: a convolution (no kernel normalisation; a registered
 * filter must be ready) admitted from the head of the queue takes along queued
 * requests of the same kind, n and flags from the next 256 queue entries (at most 16
 * requests, 2^24 values; single FP32 rows, which run through transforms of n/2
 * values, only with single rows), which are thereby admitted ahead of requests
 * between them; the group runs as one batch and its members complete
 * together.  Each member keeps its id, output and release.  Results are bit-identical
 * to the same request served alone, grouped or not. */
#define ESTL_REQ_QUEUED 0
#define ESTL_REQ_RUNNING 1
#define ESTL_REQ_DONE 2
#define ESTL_REQ_FAILED 3

int estl_server_create(estl_server_t *srv, long long hbm_budget_bytes, int streams);

/* Batched complex FFT of `batch` interleaved rows of n (FFT_OpenCL_Contiguous.h:97-142
 * semantics); the output (batch * n complex values) lives in the arena until release. */
int estl_server_submit_fft(estl_server_t srv, long n, long batch, int direction, const void *d_in,
                           void *input_stream, long long *request_id);

/* Circular convolution of `rows` real rows of length n with the k-value filter d_h
 * (estl_conv_* semantics, Convolution.h:95-148); output rows * n floats in the arena. */
int estl_server_submit_conv(estl_server_t srv, long n, long rows, const float *d_x, const float *d_h, long k,
                            unsigned conv_flags, void *input_stream, long long *request_id);

/* Register a filter for later requests: its spectrum for rows of length n (with the
 * ESTL_CONV_CENTER / ESTL_CONV_NORMALIZE_KERNEL options) is computed once and stays in
 * the arena, counted against the budget, until estl_server_release(filter_id).  The
 * registration is itself a request (wait on it or not; it is not reported by
 * estl_server_next_done).  The role of the kernel transform that
 * Convolution2D_FFT::convolve redoes on every call (Convolution.h:104-123). */
int estl_server_register_filter(estl_server_t srv, long n, const float *d_h, long k, unsigned conv_flags,
                                void *input_stream, long long *filter_id);

/* Convolution of `rows` real rows of the filter's n with a registered filter;
 * conv_flags may hold ESTL_CONV_CORRELATE only (the filter must not be FP16). */
int estl_server_submit_conv_filter(estl_server_t srv, long long filter_id, long rows, const float *d_x,
                                   unsigned conv_flags, void *input_stream, long long *request_id);

/* FP16 storage (see estl_conv_execute_fp16): d_x16 holds rows * n binary16 reals with
 * per-row scales d_x_scale; the output (in the arena) is rows * n binary16 reals followed,
 * at the next 256-byte boundary, by rows FP32 output scales.  A registered filter used
 * by FP16 requests is registered with ESTL_CONV_FP16 in its flags. */
int estl_server_submit_conv_fp16(estl_server_t srv, long n, long rows, const void *d_x16, const float *d_x_scale,
                                 const float *d_h, long k, unsigned conv_flags, void *input_stream,
                                 long long *request_id);
int estl_server_submit_conv_filter_fp16(estl_server_t srv, long long filter_id, long rows, const void *d_x16,
                                        const float *d_x_scale, unsigned conv_flags, void *input_stream,
                                        long long *request_id);

/* State (ESTL_REQ_*), the output's device address once DONE, latency submit->done. */
int estl_server_poll(estl_server_t srv, long long request_id, int *state, void **d_out, double *latency_s);

/* Block until the request is DONE (ESTL_OK), FAILED (its error) or the timeout passes
 * (ESTL_ERR_BUSY). */
int estl_server_wait(estl_server_t srv, long long request_id, double timeout_s);

/* Completion queue: up to max_ids ids of requests that finished (DONE or FAILED), oldest
 * first, waiting up to timeout_s for the first; returns how many (0 on timeout). */
int estl_server_next_done(estl_server_t srv, long long *request_ids, int max_ids, double timeout_s);

/* Copy `bytes` of a DONE request's output to d_dst on `stream` (asynchronous).  The
 * request may be released right after: its output block returns to the arena only once
 * the copies queued here have run. */
int estl_server_copy_output(estl_server_t srv, long long request_id, void *d_dst, long long bytes, void *stream);

/* Return the request's output to the arena (after DONE or FAILED).  A FAILED request
 * whose launches were partly queued has already waited for them (it turns FAILED only
 * then), so its memory is never reused under a launch still running. */
int estl_server_release(estl_server_t srv, long long request_id);

/* Wait until nothing is queued or running. */
int estl_server_drain(estl_server_t srv, double timeout_s);

/* JSON statistics (counts, p50/p99/mean/max latency in ms, throughput, arena use,
 * request groups formed and the requests they served); returns the full length. */
int estl_server_stats(estl_server_t srv, char *json, int len);
int estl_server_reset_stats(estl_server_t srv);

int estl_server_destroy(estl_server_t srv);

#ifdef __cplusplus
}
#endif

#endif /* ESTL_HIP_H */
