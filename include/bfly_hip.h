/*
 * bfly_hip.h -- C ABI of libbfly_hip.so, the gfx950 (MI355X) FFT library.
 *
 * Plain C: pointers, sizes and int status codes only; no C++ or HIP types
 * cross this boundary (a stream is an opaque void*, a hipStream_t underneath;
 * NULL means the default stream).  The C++ bfly::Context / bfly::Plan API
 * (include/butterfly/plan.hpp) is built on these functions and nothing else,
 * and the same symbol set is implemented by a CPU stub for host-only tests.
 *
 * Data layouts are the reference's transform contract
 * (docs/user-guide/transforms.md:15-44 of the reference):
 *   C2C  interleaved complex  [batch][2n] floats  -> [batch][2n]
 *   R2C  real                 [batch][n]          -> half spectrum [batch][n+2]
 *   C2R  half spectrum        [batch][n+2]        -> real [batch][n], scaled 1/n
 * Transforms are unnormalised except C2R.  Device pointers throughout the
 * execute calls; in and out must not overlap.
 *
 * Status codes map 1:1 onto the reference's Plan::run* return codes
 * (docs/reference/api.md:38-48, src/framework/butterfly.cpp:723-850).
 */
#ifndef BFLY_HIP_H
#define BFLY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define BFLY_API __attribute__((visibility("default")))
#else
#define BFLY_API
#endif

#define BFLY_OK 0
#define BFLY_ERR_NOT_LOADED (-1)  /* kernel for this transform not available   */
#define BFLY_ERR_TABLE (-2)       /* table preparation failed                  */
#define BFLY_ERR_ALLOC (-3)       /* device allocation failed                  */
#define BFLY_ERR_H2D (-4)         /* host-to-device copy failed                */
#define BFLY_ERR_LAUNCH (-5)      /* kernel launch failed                      */
#define BFLY_ERR_SYNC (-6)        /* stream synchronisation failed             */
#define BFLY_ERR_D2H (-7)         /* device-to-host copy failed                */
#define BFLY_ERR_SHAPE (-8)       /* shape / pointer / size guard              */
#define BFLY_ERR_SCRATCH (-9)     /* bfly_c2c_noalloc: the shape's intermediate
                                     exceeds BFLY_NOALLOC_SCRATCH_BYTES       */

#define BFLY_KIND_C2C 0
#define BFLY_KIND_R2C 1
#define BFLY_KIND_C2R 2

/* ------------------------------------------------------------- runtime
 * Replaces the AscendCL runtime calls of the reference's Plan/Context:
 * aclInit + aclrtSetDevice (src/framework/butterfly.cpp:574-575),
 * aclrtCreateStream/DestroyStream (butterfly.cpp:576, 556),
 * aclrtMalloc/aclrtFree (butterfly.cpp:257-260, 303-310),
 * aclrtMemcpy H2D/D2H (butterfly.cpp:263-269, 355, 368),
 * aclrtSynchronizeStream (butterfly.cpp:367). */
BFLY_API int bfly_rt_init(int device);
/* Waits for all work the library queued on the current device; the library
 * keeps no device allocation or stream of its own to release (the teardown
 * half of aclFinalize / aclrtResetDevice in the reference's Context,
 * src/framework/butterfly.cpp:552-561, without resetting a device the
 * caller may share). */
BFLY_API int bfly_rt_teardown(void);
BFLY_API int bfly_rt_stream_create(void** stream);
BFLY_API int bfly_rt_stream_destroy(void* stream);
BFLY_API int bfly_rt_stream_sync(void* stream);
BFLY_API int bfly_rt_malloc(void** ptr, size_t bytes);
/* Free and total device memory of the current device, bytes (the HBM query
 * the reference's stress runner makes through npu-smi,
 * scripts/run_stress.py:151-169). */
BFLY_API int bfly_rt_mem_info(size_t* free_bytes, size_t* total_bytes);
BFLY_API int bfly_rt_free(void* ptr);
BFLY_API int bfly_rt_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
BFLY_API int bfly_rt_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
/* The same two copies without the stream synchronisation (aclrtMemcpyAsync
 * of the reference's long-FFT device chain and end-to-end loop,
 * src/host/fft_check.cpp:853, 877, 1043-1047). */
BFLY_API int bfly_rt_memcpy_h2d_async(void* dst, const void* src, size_t bytes, void* stream);
BFLY_API int bfly_rt_memcpy_d2h_async(void* dst, const void* src, size_t bytes, void* stream);
/* Page-locked host memory (aclrtMallocHost / aclrtFreeHost of the
 * end-to-end loop, src/host/fft_check.cpp:1023-1034). */
BFLY_API int bfly_rt_malloc_host(void** ptr, size_t bytes);
BFLY_API int bfly_rt_free_host(void* ptr);
/* Stream events (aclrtCreateEvent / aclrtRecordEvent / aclrtEventElapsedTime
 * of the device-chain timing, src/host/fft_check.cpp:823-830, 852-896).
 * bfly_rt_event_elapsed waits for `end`, then gives end - start in ms. */
BFLY_API int bfly_rt_event_create(void** event);
BFLY_API int bfly_rt_event_destroy(void* event);
BFLY_API int bfly_rt_event_record(void* event, void* stream);
BFLY_API int bfly_rt_event_elapsed(void* start, void* end, float* ms);
/* Elapsed milliseconds of `reps` back-to-back launches of one transform on
 * `stream`, measured with HIP events (replaces the host wall-clock loop of
 * Plan::measure, butterfly.cpp:852-924). */
BFLY_API int bfly_rt_time_transform(int kind, int sign, const float* in, float* out, uint32_t n,
                           uint32_t batch, const float* table, void* work, int reps,
                           void* stream, float* ms);

/* Host-to-host transform of host buffers, pipelined over the PCIe link: the
 * batch is cut into chunks of `chunk_rows` transforms (0: about 8 MiB of
 * input per chunk) and chunk k's upload (on `stream`), transform and
 * download (on two streams of the library's own, created once per device by
 * bfly_rt_init or the first call) are ordered by events, so uploads,
 * transforms and downloads of different chunks overlap and the full-duplex
 * link carries both directions at once.  A batch of one chunk runs its
 * three steps on `stream` alone.
 * Replaces the synchronous upload -> kernel -> download of the reference's
 * Plan::run / runR2C / runC2R (src/framework/butterfly.cpp:723-850) for
 * callers holding host data; returns once the result is in `host_out`.
 *   kind BFLY_KIND_C2C: sign -1 / +1, unnormalised, rows of 2n floats;
 *   BFLY_KIND_R2C: [batch][n] -> [batch][n+2]; BFLY_KIND_C2R:
 *   [batch][n+2] -> [batch][n] scaled 1/n (full spectra; sign ignored).
 * Every chunk runs one kernel variant: `variant` as for bfly_c2c_ex /
 * bfly_real_ex, BFLY_VARIANT_AUTO meaning the AUTO choice for a chunk's
 * batch, so the result equals bfly_c2c_ex / bfly_real_ex of that variant on
 * the whole batch bit for bit.  `dev_in` / `dev_out` hold the whole batch
 * on the device; `work` is the long-C2C workspace of one chunk, 8 n (chunk
 * rows) bytes, where the chunk's kernel needs one (n > 16384, or the
 * four-step chain), else NULL.  Host memory may be pinned
 * (bfly_rt_malloc_host) or pageable; pageable buffers (the runtime stages
 * their copies itself, one direction at a time) run the three steps in turn
 * on `stream`: the whole batch at once, or chunk by chunk where the kernel
 * needs the workspace.  Calls on one device serialise on the library's
 * streams.  Status codes as for the other entry points
 * (BFLY_ERR_H2D / _LAUNCH / _D2H / _SYNC / _SHAPE); after an error every
 * issued step has finished before the call returns. */
BFLY_API int bfly_host_pipelined(int kind, int sign, int variant, const float* host_in, float* host_out,
                                 uint32_t n, uint32_t batch, const float* table, void* dev_in, void* dev_out,
                                 void* work, uint32_t chunk_rows, void* stream);

/* ------------------------------------------------------------- hipGraph
 * Capture of the device-pointer calls issued on `stream` into a replayable
 * graph: bfly_graph_begin starts capturing the stream (thread-local mode;
 * nothing executes while capturing), bfly_graph_end ends the capture and
 * instantiates it into *graph, bfly_graph_launch replays it on a stream and
 * bfly_graph_destroy frees it.  Capturable: bfly_c2c / bfly_c2c_ex /
 * bfly_c2c_scaled / bfly_c2c_rows, bfly_r2c / bfly_c2r / their _ex forms,
 * bfly_real_ex, bfly_fft2d, bfly_c2c_columns, bfly_mode_mix,
 * bfly_group_launch and the _async copies;
 * not the synchronising helpers (bfly_rt_memcpy_h2d / _d2h,
 * bfly_rt_stream_sync, bfly_rt_time_transform, bfly_rt_event_elapsed).
 * `stream` must be a created stream, not the null stream.  No counterpart in
 * the reference, whose Plan::run is synchronous (src/framework/butterfly.cpp:723-749);
 * the thin C ABI's graph helpers of its Context/Plan replacement. */
BFLY_API int bfly_graph_begin(void* stream);
BFLY_API int bfly_graph_end(void* stream, void** graph);
BFLY_API int bfly_graph_launch(void* graph, void* stream);
BFLY_API int bfly_graph_destroy(void* graph);

/* ---------------------------------------------------------- shape rules */
/* 1 when the library has kernels for (kind, n), else 0.  C2C: n a power of
 * two in [16, 65536]; R2C/C2R: n a power of two in [32, 8192]. */
BFLY_API int bfly_supported(int kind, uint32_t n);

/* ------------------------------------------------------- plan tables
 * The input-independent tables one transform needs (FP64-computed twiddles
 * stored as fp32), generated on the host and uploaded once per length.
 * Replaces Generator::genTwiddles + genBitReverse* + genTransposePlane +
 * genInterleave and genR2CTwiddles (butterfly.cpp:96-166, 692-703) and the
 * long-FFT wT table (src/host/fft_check.cpp:524-536): the gfx950 kernels use
 * Stockham autosort passes, so no permutation tables exist. */
BFLY_API size_t bfly_table_floats(int kind, uint32_t n);
BFLY_API int bfly_table_fill(int kind, uint32_t n, float* host_table);

/* Device workspace (bytes) an execute call with BFLY_VARIANT_AUTO needs; 0
 * for every shape that runs in one kernel (C2C up to 8192, 16384 with
 * batch >= 256; all real transforms).  The four-step path (C2C above 16384,
 * 16384 with batch < 256, and BFLY_VARIANT_FOUR_STEP) needs n*batch complex
 * values for the intermediate. */
BFLY_API size_t bfly_workspace_bytes(int kind, uint32_t n, uint32_t batch);

/* --------------------------------------------------------------- execute
 * Asynchronous on `stream`; device pointers.  `table` is the device copy of
 * bfly_table_fill(kind, n); `work` is a device buffer of at least
 * bfly_workspace_bytes(kind, n, batch) bytes (may be NULL when that is 0).
 *
 * bfly_c2c replaces the kfft_fwd launch of Plan::run (butterfly.cpp:723-749,
 * kernel ABI butterfly.cpp:663-679) and, for n > 4096, the long-FFT launch
 * chain of src/host/fft_check.cpp:846-970.  sign = -1 forward, +1
 * unnormalised inverse (the reference's twiddle-sign flip + xflip,
 * butterfly.cpp:625, src/ascendc/fft_radix2.cpp:64-69). */
BFLY_API int bfly_c2c(const float* in, float* out, uint32_t n, uint32_t batch, int sign,
             const float* table, void* work, void* stream);
/* bfly_c2c with sign -1 (forward) and +1 (unnormalised inverse). */
BFLY_API int bfly_c2c_fwd(const float* in, float* out, uint32_t n, uint32_t batch, const float* table, void* work,
                          void* stream);
BFLY_API int bfly_c2c_inv(const float* in, float* out, uint32_t n, uint32_t batch, const float* table, void* work,
                          void* stream);
/* bfly_c2c with every output multiplied by `scale` in the transform's last
 * store (no extra pass over HBM): scale 1/n on the inverse is the normalised
 * inverse, 1/sqrt(n) both ways the unitary transform.  The reference has no
 * counterpart (its C2C is unnormalised and only C2R applies 1/n,
 * src/ascendc/fft_real.cpp:167-267); it serves callers of the public inverse
 * who want numpy's norm="backward" / "ortho" / "forward" conventions.
 * Runs the LDS-twiddle streaming kernels up to n = 8192 and AUTO's choice
 * above, so the result equals fl(scale * bfly_c2c_ex(..., variant)) bit for
 * bit for that variant.  `scale` must be finite (else BFLY_ERR_SHAPE);
 * `work` as for bfly_c2c. */
BFLY_API int bfly_c2c_scaled(const float* in, float* out, uint32_t n, uint32_t batch, int sign, float scale,
                             const float* table, void* work, void* stream);

/* bfly_c2c for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): same
 * arguments minus `work`, same kernels, same results bit for bit.  No
 * counterpart in the reference.  The four-step chain's intermediate (the
 * only scratch a C2C uses: n > 16384, and 16384 with batch < 256) lives in a
 * static device array of BFLY_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; a shape whose intermediate (8 n batch bytes) does not fit
 * returns BFLY_ERR_SCRATCH (at 64 MiB: n * batch <= 8388608, e.g. 65536 x
 * 128).  The call allocates nothing and frees nothing.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams or
 * host threads) race.  bfly_rt_init looks the array up for its device;
 * bfly_noalloc_prepare does it for the current device (otherwise the first
 * call that needs it does). */
#define BFLY_NOALLOC_SCRATCH_BYTES (64u << 20)
BFLY_API int bfly_noalloc_prepare(void);
BFLY_API int bfly_c2c_noalloc(const float* in, float* out, uint32_t n, uint32_t batch, int sign,
                              const float* table, void* stream);

/* ------------------------------------------------------------------ 2-D
 * 2-D C2C of `batch` row-major [n1][n2] arrays of interleaved complex
 * values: X[k1][k2] = sum x[j1][j2] e^{sign 2 pi i (j1 k1 / n1 + j2 k2 / n2)}
 * (sign -1 forward, +1 unnormalised inverse); n1, n2 powers of two in
 * [16, 4096].  No counterpart in the reference, which lists rectangular 2-D
 * transforms (fft2d, range-Doppler) as future work
 * (config/test_matrix.json:94-99).  `table_rows` is the device copy of
 * bfly_table_fill(BFLY_KIND_C2C, n2), `table_cols` of bfly_fft2d_table_fill(n1);
 * `work` holds bfly_fft2d_workspace_bytes(n1, n2, batch) bytes.  Two HBM
 * round trips up to n1 = 2048, three at 4096. */
/* Strided 1-D C2C along the first axis of `batch` row-major [n][ncols]
 * arrays (element stride ncols, adjacent transforms): X[k][c] = sum_j
 * x[j][c] e^{sign 2 pi i j k / n}; n a power of two in [16, 4096], any
 * ncols >= 1.  The column pass of bfly_fft2d on its own: one HBM round trip
 * up to n = 2048, two at 4096 through `work`
 * (bfly_c2c_columns_workspace_bytes; NULL allowed when that is 0).
 * `table_cols` is the device copy of bfly_fft2d_table_fill(n).  No
 * counterpart in the reference (non-contiguous layouts are future work
 * there, config/test_matrix.json:94-99, 106). */
BFLY_API size_t bfly_c2c_columns_workspace_bytes(uint32_t n, uint32_t ncols, uint32_t batch);
BFLY_API int bfly_c2c_columns(const float* in, float* out, uint32_t n, uint32_t ncols, uint32_t batch, int sign,
                              const float* table_cols, void* work, void* stream);
/* Batched 1-D C2C over padded rows: transform r reads n complex values at
 * in + 2 r idist floats and writes n at out + 2 r odist floats (idist, odist
 * >= n complex values; element stride 1).  The gaps between rows are neither
 * read nor written.  n a power of two in 16..65536, `table` the device copy
 * of bfly_table_fill(BFLY_KIND_C2C, n); in and out 8-byte aligned, their
 * spans [p, p + 8 ((batch-1) dist + n)) disjoint.  Up to 16384 runs the
 * LDS-twiddle streaming kernel (no workspace; with idist = odist = n equal
 * to bfly_c2c_ex(..., BFLY_VARIANT_STREAM_LT, ...) bit for bit); 32768 and
 * 65536 run the four-step chain, whose column pass reads the padded input
 * and row pass writes the padded output, through `work` of
 * bfly_workspace_bytes(BFLY_KIND_C2C, n, batch) bytes (with idist = odist
 * = n equal to bfly_c2c bit for bit).  No counterpart
 * in the reference: non-contiguous contracts are future work there
 * (config/test_matrix.json:94-99, 106 "fft2d-and-strides"). */
BFLY_API int bfly_c2c_rows(const float* in, uint64_t idist, float* out, uint64_t odist, uint32_t n, uint32_t batch,
                           int sign, const float* table, void* work, void* stream);
BFLY_API size_t bfly_fft2d_table_floats(uint32_t n1);
BFLY_API int bfly_fft2d_table_fill(uint32_t n1, float* host_table);
BFLY_API size_t bfly_fft2d_workspace_bytes(uint32_t n1, uint32_t n2, uint32_t batch);
BFLY_API int bfly_fft2d(const float* in, float* out, uint32_t n1, uint32_t n2, uint32_t batch, int sign,
                        const float* table_rows, const float* table_cols, void* work, void* stream);

/* Kernel launches bfly_c2c_ex makes for (n, batch, variant): 1, or 2 for
 * the four-step chain (the column kernel -- transpose in, N1-point FFTs,
 * twiddle -- then the row kernel -- N2-point FFTs, natural-order store);
 * BFLY_ERR_SHAPE for an unsupported length, batch 0 or an unknown variant. */
BFLY_API int bfly_c2c_launches(uint32_t n, uint32_t batch, int variant);
/* bfly_c2c_ex with its launches bracketed by events, for per-segment timing
 * (the evSeg markers of the reference's device chain,
 * src/host/fft_check.cpp:852-896): events[0] is recorded before the first
 * launch and events[i] after launch i, so nevents must be
 * bfly_c2c_launches(n, batch, variant) + 1. */
BFLY_API int bfly_c2c_traced(const float* in, float* out, uint32_t n, uint32_t batch, int sign,
                             int variant, const float* table, void* work, void* stream,
                             void* const* events, int nevents);

/* Kernel variants of one C2C length (same table, same results):
 *   BFLY_VARIANT_AUTO    the library's choice for (n, batch)
 *   BFLY_VARIANT_PLAIN   one workgroup per group of transforms, one pass over
 *                        the batch (grid = batch / transforms-per-workgroup)
 *   BFLY_VARIANT_STREAM  grid sized to the resident capacity; each workgroup
 *                        walks the batch with register-resident twiddles and
 *                        prefetches its next transform during the current one
 *   BFLY_VARIANT_PLAIN_SWZ / BFLY_VARIANT_STREAM_SWZ  the same two with an
 *                        XOR-swizzled, unpadded LDS exchange (32 KiB per
 *                        256-thread workgroup instead of 34 KiB)
 *   BFLY_VARIANT_STREAM_LT  streaming form with two 64-entry twiddle tables in
 *                        LDS (W = A[m mod 64] * B[m / 64]) instead of
 *                        register-resident twiddles
 *   BFLY_VARIANT_STREAM2 streaming form with two transforms in flight per
 *                        workgroup (swizzled LDS, register twiddles)
 *   BFLY_VARIANT_FOUR_STEP  n = 8192 / 16384 through the two-kernel four-step
 *                        path (needs n*batch*8 bytes of workspace) instead of
 *                        one 512 / 1024-thread workgroup per transform
 *                        (PLAIN: one-shot grid; STREAM_LT: resident
 *                        streaming grid with LDS twiddles)
 *   BFLY_VARIANT_SPLIT_MFMA n = 1024 / 2048 / 4096 only (else BFLY_ERR_SHAPE):
 *                        the radix-16 passes as 16x16 complex DFT products on
 *                        the matrix cores in split FP16 (hi + lo operands,
 *                        FP32 accumulation, per-butterfly power-of-two
 *                        scaling); FP32 twiddles.  Relative L2 error vs FP64
 *                        about 3x that of the FP32 variants, within 1e-5
 *   BFLY_VARIANT_SPLIT_MFMA_STREAM  the same on a resident grid that
 *                        prefetches each workgroup's next transforms
 * For n <= 4096 FOUR_STEP means PLAIN.  At 8192 / 16384, PLAIN and
 * STREAM_LT run one workgroup per transform (one-shot / streaming) and the
 * other ids the one-shot form; above 16384 every id runs four-step. */
#define BFLY_VARIANT_AUTO 0
#define BFLY_VARIANT_PLAIN 1
#define BFLY_VARIANT_STREAM 2
#define BFLY_VARIANT_PLAIN_SWZ 3
#define BFLY_VARIANT_STREAM_SWZ 4
#define BFLY_VARIANT_STREAM_LT 5
#define BFLY_VARIANT_STREAM2 6
#define BFLY_VARIANT_FOUR_STEP 7
#define BFLY_VARIANT_SPLIT_MFMA 8
#define BFLY_VARIANT_SPLIT_MFMA_STREAM 9
#define BFLY_VARIANT_LAST 9
BFLY_API int bfly_c2c_ex(const float* in, float* out, uint32_t n, uint32_t batch, int sign,
                         int variant, const float* table, void* work, void* stream);

/* bfly_r2c replaces kfft_fwd(n/2) + kfft_r2c_post + the host repack of
 * Plan::runR2C (butterfly.cpp:753-806); one fused kernel, output written
 * directly at row stride n+2. */
BFLY_API int bfly_r2c(const float* in, float* out, uint32_t n, uint32_t batch, const float* table,
             void* stream);

/* bfly_c2r replaces the host pad + kfft_c2r_prep + kfft_fwd(xflip) +
 * kfft_c2r_post chain of Plan::runC2R (butterfly.cpp:810-850); one fused
 * kernel.  The imaginary parts of the DC and Nyquist bins are ignored, as the
 * reference's Hermitian expansion does (src/framework/reference.cpp:64-79). */
BFLY_API int bfly_c2r(const float* in, float* out, uint32_t n, uint32_t batch, const float* table,
             void* stream);

/* Mode-truncated real transforms (the spectral-convolution layer keeps the
 * first `bins` modes): bfly_r2c_ex writes bins 0..bins-1 of each row at a row
 * stride of out_stride floats; bfly_c2r_ex reads bins 0..bins-1 at a row
 * stride of in_stride floats and treats every higher bin as zero.
 * 1 <= bins <= n/2+1; the stride is even and >= 2*bins.  bfly_r2c / bfly_c2r
 * are these with bins = n/2+1 and stride = n+2. */
BFLY_API int bfly_r2c_ex(const float* in, float* out, uint32_t n, uint32_t batch, uint32_t bins,
                         uint32_t out_stride, const float* table, void* stream);
BFLY_API int bfly_c2r_ex(const float* in, float* out, uint32_t n, uint32_t batch, uint32_t bins,
                         uint32_t in_stride, const float* table, void* stream);

/* Real transforms with an explicit kernel variant (kind R2C or C2R; bins and
 * stride as for bfly_r2c_ex / bfly_c2r_ex): BFLY_VARIANT_PLAIN = one-shot
 * grid, BFLY_VARIANT_STREAM = resident grid that loads the next rows while
 * computing the current ones (same arithmetic, same bits),
 * BFLY_VARIANT_STREAM_LT = streaming with the twiddles in LDS (two-level
 * products; n >= 256, else STREAM), BFLY_VARIANT_AUTO = the library's
 * choice; other ids give BFLY_ERR_SHAPE.  Same reference interfaces as
 * bfly_r2c / bfly_c2r.
 *
 * flags: 0, or BFLY_REAL_ADJOINT for the transform's adjoint in the same
 * layouts (the backward pass of a differentiable FFT, one launch, no extra
 * weighting pass):
 *   R2C + ADJOINT: out[k] = X[k] * v_k / n, v_k = 1 at bins 0 and n/2 (with
 *     the imaginary part set to 0) and 2 elsewhere -- the adjoint of C2R;
 *   C2R + ADJOINT: out = n * C2R(w X), w_k = 1 at bins 0 and n/2 and 1/2
 *     elsewhere -- the adjoint of R2C.
 * Every weight is a power of two: the results equal the forward transforms
 * followed by the weighting, bit for bit.  Other flag bits give
 * BFLY_ERR_SHAPE.  The reference has no adjoint transform. */
#define BFLY_REAL_ADJOINT 1u
BFLY_API int bfly_real_ex(int kind, const float* in, float* out, uint32_t n, uint32_t batch,
                          uint32_t bins, uint32_t stride, int variant, uint32_t flags,
                          const float* table, void* stream);

/* ------------------------------------------------------- grouped launch
 * Many independent transforms of different kinds, lengths and batch sizes
 * run by ONE persistent kernel driven by a device-resident request table.
 * No counterpart in the reference: its Plan::run executes one (n, batch)
 * shape per launch (butterfly.cpp:723-749), so a caller with many uneven
 * requests pays one launch and one tail per request.  The table is built and
 * uploaded once by bfly_group_create; bfly_group_launch only launches, so it
 * can be captured into a hipGraph and replayed (the table's device address
 * stays fixed until bfly_group_destroy).  Results are those of per-request
 * bfly_c2c_ex / bfly_real_ex calls of variant PLAIN (STREAM_LT with
 * twiddles = 1): with schedule 0 the same kernels' per-transform code runs,
 * bit for bit.
 *
 * Work is cut into units of 16*workgroup points (one HBM read and write
 * each): a unit is workgroup/(n'/16) transforms of one request, n' = n for
 * C2C and n/2 for the real kinds.  Lengths: C2C 16..8192 (16..16384 with
 * 1024-thread workgroups), R2C/C2R 32..8192. */
typedef struct bfly_group_request {
    const void* in;   /* device input in the kind's execute layout */
    void* out;        /* device output in the kind's execute layout */
    int32_t kind;     /* BFLY_KIND_C2C / R2C / C2R */
    int32_t sign;     /* C2C: -1 forward, +1 unnormalised inverse; else ignored */
    uint32_t n;       /* transform length */
    uint32_t batch;   /* transforms, >= 1 */
} bfly_group_request;

typedef struct bfly_group_config {
    int32_t workgroup; /* threads per workgroup: 512 or 1024; 0 = 512 */
    int32_t schedule;  /* radix schedule: 0 = radix 16 first, remainder radix last;
                          1 = remainder radix first */
    int32_t order;     /* unit order: 0 = request order; 1 = most work per unit
                          first; 2 = round robin over requests; 3 = requests
                          with the least total work first (lowest median
                          completion time) */
    int32_t dynamic;   /* 0 = workgroup w runs units w, w+grid, ...;
                          1 = workgroups take units from an atomic counter */
    int32_t grid;      /* workgroups launched; 0 = the resident capacity */
    int32_t twiddles;  /* 0 = inter-pass twiddles read from device memory (the
                          per-request PLAIN kernels' code); 1 = each unit stages
                          its plan's two-level tables in LDS first (the
                          per-request BFLY_VARIANT_STREAM_LT code, bit for bit;
                          schedule 0 only; real kinds below n = 256 keep the
                          device tables, as their STREAM_LT launches do) */
} bfly_group_config;

/* 1 when a grouped launch with `workgroup` threads (512 / 1024) takes
 * (kind, n), else 0. */
BFLY_API int bfly_group_supported(int kind, uint32_t n, int workgroup);
/* Validates the requests, builds the request table, the unit list and the
 * twiddle tables of every distinct length on the host and uploads them into
 * one device allocation (hipMalloc + synchronous copy; call outside stream
 * capture).  cfg may be NULL (all defaults).  *group receives the handle. */
BFLY_API int bfly_group_create(const bfly_group_request* reqs, int nreq,
                               const bfly_group_config* cfg, void** group);
/* Launches the group on `stream`; allocates nothing, copies nothing.  When
 * `stamps` is a device array of nreq+1 uint64, a small kernel first clears
 * it, then the grouped kernel records stamps[0] = earliest workgroup start
 * and stamps[1+r] = completion of request r's last unit, in ticks of the
 * device wall clock (bfly_probe_device reports its rate).  Two launches of
 * one group on different streams must not overlap (the dynamic schedule's
 * counter is per group). */
BFLY_API int bfly_group_launch(void* group, unsigned long long* stamps, void* stream);
/* units = work units, grid = workgroups launched, bytes = device bytes held. */
BFLY_API int bfly_group_query(void* group, uint32_t* units, uint32_t* grid, size_t* bytes);
BFLY_API int bfly_group_destroy(void* group);

/* ------------------------------------------------- spectral convolution
 * Per-mode channel mixing of the 1-D spectral-convolution layer, in the
 * layouts bfly_r2c_ex / bfly_c2r_ex use for truncated modes (complex64
 * interleaved, mode index fastest):
 *   x  [batch][c_in][modes]   input modes        (bfly_r2c_ex, bins = modes)
 *   w  [c_in][c_out][modes]   weights
 *   y  [batch][c_out][modes]  output modes       (bfly_c2r_ex, bins = modes)
 * op BFLY_MIX_FORWARD:      a = x,  b = w,  c = y,   y[b][o][k] = sum_i x[b][i][k] w[i][o][k]
 * op BFLY_MIX_GRAD_INPUT:   a = gy, b = w,  c = gx,  gx[b][i][k] = sum_o gy[b][o][k] conj(w[i][o][k])
 * op BFLY_MIX_GRAD_WEIGHT:  a = x,  b = gy, c = gw,  gw[i][o][k] = sum_b conj(x[b][i][k]) gy[b][o][k]
 * (the gradients follow torch's convention for complex tensors).  Every
 * element of c is written; c must not overlap a or b.  fp32 multiply-adds,
 * fp32 accumulation.  No counterpart in the reference (which has no
 * spectral layer); it replaces the transposed copies plus batched GEMM a
 * torch implementation of the mixing needs. */
#define BFLY_MIX_FORWARD 0
#define BFLY_MIX_GRAD_INPUT 1
#define BFLY_MIX_GRAD_WEIGHT 2
BFLY_API int bfly_mode_mix(int op, const float* a, const float* b, float* c, uint32_t batch, uint32_t c_in,
                           uint32_t c_out, uint32_t modes, void* stream);

/* ------------------------------------------------------------- probes */
/* Device-to-device copy of `bytes` (16-byte aligned, multiple of 16) with
 * the library's streaming copy kernel: the measured HBM ceiling next to the
 * nominal peak (replaces src/host/bw_probe.cpp). */
BFLY_API int bfly_probe_copy(const void* src, void* dst, size_t bytes, void* stream);

/* Device facts for the gfx950 hardware profile (replaces
 * src/host/profile.cpp).  Writes at most `cap` bytes of JSON into `json`. */
BFLY_API int bfly_probe_device(int device, char* json, size_t cap);

#ifdef __cplusplus
}
#endif
#endif /* BFLY_HIP_H */
