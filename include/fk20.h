/*
 * fk20.h -- C ABI of the MI355X (gfx950) FK20 multi-proof library, libfk20.so.
 *
 * Drop-in replacement for the per-row FK20 path of lCardosoSantos/FK20-CUDA
 * (fk20.cuh:20-38, fr.cuh:62-69, g1.cuh:83-100).  The reference exposes CUDA kernels
 * with a launch-shape contract (grid = rows, 256 threads, dynamic shared memory) and
 * void returns; here every entry point is a host function that takes device pointers,
 * the row/transform count, the size, and a stream, and returns a status code.
 *
 * Types are byte-compatible with the reference ABI:
 *   fk20_fr_t  = fr_t   (fr.cuh:24)   4 x u64 little endian, standard form
 *   fk20_g1p_t = g1p_t  (g1.cuh:27)   x, y, z, each 6 x u64, homogeneous projective
 *   fk20_g1a_t = g1a_t  (g1.cuh:20)   x, y affine, infinity = (0, 0) (g1a.cu:91)
 * Inputs may be any representative (fr: any 256-bit value, fp: any 384-bit value).
 * Outputs are standard form with every coordinate canonical (< p or < r); projective
 * outputs are not normalised (use fk20_g1p_to_affine for canonical affine points).
 * G1 inputs may be any point of E(Fp), as in the reference.  On the order-r subgroup G1 (all
 * FK20 data lie in it) a transform is a linear map mod r; on a point outside it [a]P and [b]P
 * differ for a = b (mod r), so the transforms check their inputs and recompute a row holding
 * such a point in the reference's exact order of operations (same butterflies, the canonical
 * twiddles and 2^-9 as double-and-add multipliers), giving the reference's points.
 * fk20_poly2h_fft does so from the MSM on, whose scalars are canonical here (the reference's
 * may be other representatives mod r).  The batched prover (fk20_setup_load) takes setups in
 * G1 only (FK20_EINVAL otherwise); fk20_commit_setup_load takes any point except one of small
 * order, whose multiples reach infinity inside the affine window tables (FK20_EINVAL).
 *
 * Sizes: N is the polynomial length (1024, 2048 or 4096; the reference fixes 4096),
 * l = 16 is fixed, k = N/16, and n2 = 2k = N/8 is the FFT size (128, 256, 512).
 * All buffers are device memory; aliasing of input and output is allowed where the
 * reference allows it (fr_fft.cu:21, g1p_fft.cu:17).  Calls are asynchronous on
 * `stream` (a hipStream_t; NULL = the default stream) unless stated otherwise.
 */
#ifndef FK20_H
#define FK20_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint64_t v[4]; } fk20_fr_t;
typedef struct { uint64_t x[6], y[6], z[6]; } fk20_g1p_t;
typedef struct { uint64_t x[6], y[6]; } fk20_g1a_t;
typedef void *fk20_stream_t; /* hipStream_t */

enum {
    FK20_OK = 0,
    FK20_EINVAL = -1,   /* bad size, count or null pointer */
    FK20_ENOMEM = -2,   /* device allocation failed or budget exceeded */
    FK20_EHIP = -3,     /* HIP runtime error (fk20_last_hip_error gives the code) */
    FK20_ENODEV = -4,   /* no usable gfx950 device */
    FK20_ETIMEDOUT = -5, /* a bounded device-side wait expired; the outputs are not valid */
    FK20_ERANGE = -6     /* request larger than the fixed workspace of fk20_poly2h_fft_noalloc */
};

/* Last HIP error code seen by the library in this thread (0 if none). */
int fk20_last_hip_error(void);
const char *fk20_status_string(int status);

/* ------------------------------------------------------------------ drop-in entry points */

/* polynomial[rows][N] -> toeplitz_coefficients[rows][16][n2].
 * Replaces fk20_poly2toeplitz_coefficients (fk20.cuh:25; fk20_poly2toeplitz_coefficients.cu:25). */
int fk20_poly2toeplitz_coefficients(fk20_fr_t *toeplitz_coefficients, const fk20_fr_t *polynomial,
                                    unsigned rows, unsigned N, fk20_stream_t stream);

/* polynomial[rows][N] -> toeplitz_coefficients_fft[rows][16][n2] (Toeplitz + NTT fused).
 * Replaces fk20_poly2toeplitz_coefficients_fft (fk20.cuh:26; fk20_poly2toeplitz_coefficients_fft.cu:24). */
int fk20_poly2toeplitz_coefficients_fft(fk20_fr_t *toeplitz_coefficients_fft,
                                        const fk20_fr_t *polynomial, unsigned rows, unsigned N,
                                        fk20_stream_t stream);

/* count NTTs of size n (128/256/512): output[i] = sum_j input[j] w^(ij), w = 5^((r-1)/n).
 * Replaces fr_fft_wrapper (fr.cuh:67; fr_fft.cu:318) and fr_fft (fr.cuh:62). */
int fr_fft(fk20_fr_t *output, const fk20_fr_t *input, unsigned count, unsigned n,
           fk20_stream_t stream);
/* Inverse NTT including the 1/n factor.  Replaces fr_ift_wrapper (fr.cuh:68; fr_fft.cu:347). */
int fr_ift(fk20_fr_t *output, const fk20_fr_t *input, unsigned count, unsigned n,
           fk20_stream_t stream);

/* count G1 FFTs of size n.  Replaces g1p_fft_wrapper (g1.cuh:96; g1p_fft.cu:337).  The G1
 * transforms (and fk20_hext_fft2h_fft, fk20_poly2h_fft, fk20_setup2xext_fft) stage their
 * points in library-owned scratch of count * n * 144 bytes (grown on demand, per device,
 * shared by these calls: not reentrant), so output may alias input. */
int g1p_fft(fk20_g1p_t *output, const fk20_g1p_t *input, unsigned count, unsigned n,
            fk20_stream_t stream);
/* Inverse G1 FFT including 1/n.  Replaces g1p_ift_wrapper (g1.cuh:97; g1p_fft.cu:350). */
int g1p_ift(fk20_g1p_t *output, const fk20_g1p_t *input, unsigned count, unsigned n,
            fk20_stream_t stream);

/* hext_fft[rows][n2] = sum_i tc_fft[rows][i][n2] * xext_fft[i][n2] (16 terms per column).
 * Replaces fk20_msm (fk20.cuh:33; fk20_msm.cu:23). */
int fk20_msm(fk20_g1p_t *hext_fft, const fk20_fr_t *toeplitz_coefficients_fft,
             const fk20_g1p_t *xext_fft, unsigned rows, unsigned n2, fk20_stream_t stream);

/* h[rows][n2]: upper half set to infinity in place.  Replaces fk20_hext2h (fk20.cuh:32;
 * fk20_poly2h_fft.cu:27). */
int fk20_hext2h(fk20_g1p_t *h, unsigned rows, unsigned n2, fk20_stream_t stream);

/* h_fft = FFT(upper half zeroed(IFFT(hext_fft))) per row, fused.
 * Replaces fk20_hext_fft2h_fft (fk20.cuh:31; fk20_hext_fft2h_fft.cu:27). */
int fk20_hext_fft2h_fft(fk20_g1p_t *h_fft, const fk20_g1p_t *hext_fft, unsigned rows, unsigned n2,
                        fk20_stream_t stream);

/* The plain per-row launch sequence: polynomial[rows][N] + xext_fft[16][n2] -> h_fft[rows][n2].
 * Replaces fk20_poly2h_fft (fk20.cuh:22; fk20_poly2h_fft.cu:47-80).  Like the reference it
 * uses library-owned scratch (grown on demand, per device) and is not reentrant; use the
 * context API below for concurrent or allocation-free use. */
int fk20_poly2h_fft(fk20_g1p_t *h_fft, const fk20_fr_t *polynomial, const fk20_g1p_t *xext_fft,
                    unsigned rows, unsigned N, fk20_stream_t stream);

/* fk20_poly2h_fft without any allocation, for callers that cannot allocate device memory
 * (code under stream capture, tight serving loops, hosts that forbid hipMalloc).  No
 * reference counterpart: the reference's fk20_poly2h_fft owns static managed scratch
 * (fk20_poly2h_fft.cu:13-16) but has no allocation-free contract.  The workspace is a fixed
 * FK20_NOALLOC_SCRATCH_BYTES array in device memory, one per device the library runs on, held
 * until the process ends; a call needs rows * (N/8) * 1344 + (N/8) * 2304 bytes of it, so
 * rows <= 95 at N = 4096, 193 at N = 2048 and 388 at N = 1024, and larger requests return
 * FK20_ERANGE without touching the device.  Every call on a device shares that device's
 * array: calls that overlap on one device (different streams, threads) race and must be
 * serialised by the caller.  The first call of any entry point on a device copies the
 * library's constant tables into static device memory (no allocation there either). */
#define FK20_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int fk20_poly2h_fft_noalloc(fk20_g1p_t *h_fft, const fk20_fr_t *polynomial, const fk20_g1p_t *xext_fft,
                            unsigned rows, unsigned N, fk20_stream_t stream);

/* setup SRS [tau^i]G (i = 0..N) -> xext_fft[16][n2] by FK20Py toeplitz_part1.
 * Replaces fk20_setup2xext_fft (fk20.cuh:20; fk20.cu:28), whose reference version is
 * marked as failing its test (fk20.cu:29); this one follows FK20Py fk20_multi.py:82-85. */
int fk20_setup2xext_fft(fk20_g1p_t *xext_fft, const fk20_g1p_t *setup, unsigned N,
                        fk20_stream_t stream);

/* ------------------------------------------------------------------ batched path (context API)
 *
 * The reference's fast 512-row path (fk20benchmark.cu:287-309: fk20_msm_makelut once per
 * setup, then fk20_msm_comb and the graph-captured G1 FFT) as a context that keeps each
 * setup's tables resident and owns its workspace, so proving allocates nothing once the
 * workspace is reserved.  A context is bound to the device current at creation; calls on
 * one context are serialised by the library: host-side by a lock, and on the device by an
 * event, so a fk20_prove_rows / fk20_prove_batch / fk20_commit_rows call on one stream
 * starts after the previous such call on another stream has finished with the context's
 * workspace (fk20_prove_batch_ws, which brings its own workspace, is not ordered).  No
 * reference counterpart exists for the context type itself; fk20_setup_load replaces
 * fk20_msm_makelut (fk20.cuh:34;
 * fk20_msm.cu:79) and fk20_prove_rows replaces the fk20_poly2h_fft sequence with
 * fk20_msm_comb (fk20.cuh:35) and fk20_hext_fft_2_h_fft_512 (fk20.cuh:38). */
typedef struct fk20_ctx fk20_ctx_t;

int fk20_ctx_create(fk20_ctx_t **ctx);
int fk20_ctx_destroy(fk20_ctx_t *ctx);
/* 1: fk20_prove_rows captures its launch sequence into a hipGraph per (setup, rows,
 * buffers, stream) and replays it (needs a non-NULL stream); 0: direct launches. */
int fk20_ctx_set_graph(fk20_ctx_t *ctx, int enable);
/* 1: the G1 IFFT / zero / FFT stages of a batch run as one persistent launch in which each
 * 64-row butterfly starts as soon as the two butterflies it reads have finished; 0 (default):
 * one launch per stage.  Results are identical; on MI355X the per-stage launches are faster
 * (in the persistent form a wave that has taken a butterfly holds its slot while it waits for
 * the butterfly's inputs, so fewer ladders run at once).  Each of those waits is bounded;
 * with the persistent form on, fk20_prove_rows / fk20_prove_batch synchronise `stream` and
 * return FK20_ETIMEDOUT if any wait expired (its butterfly then ran on unfinished inputs).
 * fk20_prove_batch_ws always uses the per-stage launches. */
int fk20_ctx_set_chain(fk20_ctx_t *ctx, int enable);
/* Bound of each wait of the persistent form, in ticks of the device's constant-rate wall
 * clock (100 MHz on MI355X); default 2^27 (~1.3 s, where a legitimate wait lasts at most one
 * pass, tens of ms).  0 makes every wait that finds its inputs unfinished give up at once:
 * a test of the FK20_ETIMEDOUT path in which every kernel still exits normally. */
int fk20_ctx_set_chain_wait(fk20_ctx_t *ctx, uint64_t ticks);
/* 2 (default): fk20_prove_rows / fk20_prove_batch split the rows into two halves at a 64-row
 * boundary and run the halves' launch sequences on two internal streams forked from and
 * joined back to the caller's stream (captured as such in a hipGraph), so that one half's
 * stage launches fill the SIMDs the other half's leave idle at a launch boundary;
 * 1: a single sequence on the caller's stream.  Results are identical.  fk20_prove_batch_ws
 * (the serving loop's, which overlaps batches itself) always uses the caller's stream. */
int fk20_ctx_set_streams(fk20_ctx_t *ctx, unsigned n);
/* Schedule of the transform kernels, per context (the choices an autotuner sweeps; results
 * are identical for every schedule):
 *   wnaf            3, 4 or 5: digit width of the twiddle programs, i.e. a table of 2, 4 or 8
 *                   odd multiples per ladder (default 5)
 *   radix           4: radix-2^2 butterflies (only w^+-128 twiddles) with point-wise twiddle
 *                   layers (default); 2: radix-2 butterflies carrying every twiddle
 *   waves_per_simd  2 (default): kernels compiled to leave room for two waves per SIMD
 *                   (up to 256 registers each); 3: three (168 registers, the rest spilled)
 * The reference's counterparts are compile-time switches (FFT_Sharedmem, fk20benchmark.cu:302;
 * SHAREDMEMROOTS, fk20_hext_fft2h_fft.cu:4; CUDAGRAPH, fk20benchmark.cu:5).  The persistent
 * form (fk20_ctx_set_chain) always runs radix-2 stages with the schedule's width and waves. */
int fk20_ctx_set_schedule(fk20_ctx_t *ctx, unsigned wnaf, unsigned radix, unsigned waves_per_simd);
int fk20_ctx_get_schedule(fk20_ctx_t *ctx, unsigned *wnaf, unsigned *radix, unsigned *waves_per_simd);
/* Output format of fk20_prove_rows / fk20_prove_batch (h_fft buffers), per context:
 *   FK20_OUT_G1P           g1p_t[rows][n2], natural order: the reference ABI (default)
 *   FK20_OUT_G1P_DA        g1p_t, data-availability order: FK20Py list_to_reverse_bit_order
 *                          of the natural list (fk20_multi.py:107-118, kzg_proofs.py:107-118)
 *   FK20_OUT_COMPRESSED    48-byte compressed G1 per point (big-endian x; flags 0x80
 *                          compressed, 0x40 infinity, 0x20 larger y), natural order
 *   FK20_OUT_COMPRESSED_DA the same in data-availability order
 * No reference counterpart: the reference emits g1p_t only. */
#define FK20_OUT_G1P 0
#define FK20_OUT_G1P_DA 1
#define FK20_OUT_COMPRESSED 2
#define FK20_OUT_COMPRESSED_DA 3
int fk20_ctx_set_output(fk20_ctx_t *ctx, int format);
int fk20_ctx_get_output(fk20_ctx_t *ctx, int *format);
/* Bytes per output point of a format (144 or 48; 0 for an unknown format). */
size_t fk20_output_point_bytes(int format);
/* Device bytes of one setup's window tables for polynomial length N and window width w
 * (8 to 13 bits): 16 * (N/8) * ceil(256/w) * 2^(w-1) * 128 (an affine entry padded to one
 * 128-byte cache line), e.g. 80 GiB at N=4096, w=13. */
size_t fk20_setup_table_bytes(unsigned N, unsigned window);
/* Builds the resident tables of xext_fft[16][N/8] (device memory, standard form);
 * synchronises `stream`.  The returned id names the setup in fk20_prove_rows.  FK20_EINVAL
 * for a setup with a point outside G1. */
int fk20_setup_load(fk20_ctx_t *ctx, const fk20_g1p_t *xext_fft, unsigned N, unsigned window,
                    int *setup_id, fk20_stream_t stream);
int fk20_setup_unload(fk20_ctx_t *ctx, int setup_id);
/* Workspace bytes fk20_prove_rows needs for `rows` rows of length N:
 * rows * (N/8) * 656 + ceil(rows/64) * (N/8) * 4 + 512. */
size_t fk20_rows_workspace_bytes(unsigned rows, unsigned N);
/* Grows the context workspace to fit `rows` rows of length N. */
int fk20_ctx_reserve(fk20_ctx_t *ctx, unsigned rows, unsigned N);
/* polynomial[rows][N] -> h_fft[rows][N/8] for the setup's N: Toeplitz + NTT, window-table
 * MSM, column-batched G1 IFFT / zero / FFT with endomorphism-split twiddles. */
int fk20_prove_rows(fk20_ctx_t *ctx, int setup_id, fk20_g1p_t *h_fft, const fk20_fr_t *polynomial,
                    unsigned rows, fk20_stream_t stream);
/* Several row groups in one batched pass (each launch of the path covers every group):
 * groups may use different setups and polynomial lengths N (1024/2048/4096), e.g. the
 * uneven batches of a serving loop.  At most 16 groups per call. */
typedef struct {
    int setup_id;
    unsigned rows;
    const fk20_fr_t *polynomial;  /* [rows][N of the setup] */
    fk20_g1p_t *h_fft;            /* [rows][N/8] */
} fk20_batch_item_t;
int fk20_prove_batch(fk20_ctx_t *ctx, const fk20_batch_item_t *items, unsigned nitems,
                     fk20_stream_t stream);
/* fk20_prove_batch with a caller-provided device workspace of at least the sum of
 * fk20_rows_workspace_bytes over the items (no allocation; used by the serving loop). */
int fk20_prove_batch_ws(fk20_ctx_t *ctx, const fk20_batch_item_t *items, unsigned nitems,
                        void *workspace, size_t workspace_bytes, fk20_stream_t stream);
/* KZG commitments of rows (the step before the path): commitments[r] = sum_{i<N}
 * polynomial[r][i] * srs[i], as FK20Py commit_to_poly (kzg_proofs.py:159, lincomb over
 * setup[0][:len(polynomial)]).  fk20_commit_setup_load builds fixed-base window tables of the
 * N SRS points srs[0..N) (device, standard-form g1p_t; synchronises `stream`); the returned
 * id is accepted only by fk20_commit_rows, which writes one canonical standard-form g1p_t per
 * row using the context workspace.  No CUDA counterpart: the reference's commitments are
 * FK20Py-only. */
int fk20_commit_setup_load(fk20_ctx_t *ctx, const fk20_g1p_t *srs, unsigned N, unsigned window,
                           int *setup_id, fk20_stream_t stream);
size_t fk20_commit_setup_table_bytes(unsigned N, unsigned window);
size_t fk20_commit_workspace_bytes(unsigned rows, unsigned N);
int fk20_commit_rows(fk20_ctx_t *ctx, int setup_id, fk20_g1p_t *commitments, const fk20_fr_t *polynomial,
                     unsigned rows, fk20_stream_t stream);
/* N and window of a loaded setup. */
int fk20_setup_info(fk20_ctx_t *ctx, int setup_id, unsigned *N, unsigned *window);

/* ------------------------------------------------------------------ serving loop
 *
 * Continuous batching over a context (no reference counterpart: the reference proves
 * fixed batches).  Requests are host polynomial blobs (rows x N of one loaded setup) with a
 * host output buffer; best in pinned memory (fk20_host_alloc) so copies are asynchronous.
 * A worker thread keeps up to `nslots` batches in flight, one per stream; a free slot takes
 * pending requests in arrival order while they fit the free part of the device-memory budget
 * for in-flight data (one arena of budget_bytes shared by the slots) -- per row the
 * toeplitz_fft (16*N/8 Fr) and hext (N/8 G1) workspace plus the polynomial and h_fft -- and
 * at most max_batch_rows rows and 16 setups, packs them per setup into one batched pass, and
 * copies results back.  Batches are enqueued outside the server's lock, so submit and wait
 * do not block behind a launch.
 * Requests arriving while batches run are admitted as slots free up; results are the same
 * as proving each request alone.  Outputs use the context's output format as it is when the
 * server is created (host_h_fft then holds rows * N/8 points of that format); batches
 * launched after a format change fail with FK20_EINVAL. */
typedef struct fk20_server fk20_server_t;
int fk20_server_create(fk20_server_t **server, fk20_ctx_t *ctx, size_t budget_bytes, unsigned nslots,
                       unsigned max_batch_rows);
/* FK20_ENOMEM if the request alone exceeds the budget (or max_batch_rows). */
int fk20_server_submit(fk20_server_t *server, int setup_id, const fk20_fr_t *host_polynomial,
                       unsigned rows, fk20_g1p_t *host_h_fft, uint64_t *request_id);
/* Blocks until the request is done; latency = completion - submission, seconds. */
int fk20_server_wait(fk20_server_t *server, uint64_t request_id, double *latency_s);
/* Batches launched, rows completed, peak rows in flight; returns the first error seen. */
int fk20_server_stats(fk20_server_t *server, uint64_t *batches, uint64_t *rows_done,
                      uint64_t *max_inflight_rows);
/* Drains pending and in-flight requests, then frees the slots. */
int fk20_server_destroy(fk20_server_t *server);
int fk20_host_alloc(void **ptr, size_t bytes);  /* pinned host memory */
int fk20_host_free(void *ptr);

/* Per-launch timing of fk20_prove_rows (direct launches, graph replay off while enabled):
 * an event is recorded after every kernel; fk20_ctx_launch_times waits for the last call's
 * events and returns per-launch milliseconds and kinds (0 Toeplitz+NTT, 1 MSM, 2 IFFT stage,
 * 3 middle stage, 4 FFT stage, 5 output, 6 persistent transform, 7 twiddle layer of the
 * radix-2^2 transform schedule). */
int fk20_ctx_set_launch_timing(fk20_ctx_t *ctx, int enable);
int fk20_ctx_launch_times(fk20_ctx_t *ctx, float *ms, int *kinds, int max, int *count);
/* Host-only check of the twiddle programs (GLV split + wNAF of every w^+-i, i < 512,
 * evaluated over Fr); total_ops receives the program size and wnaf_width the digit width
 * the programs (and the transform kernels' tables of odd multiples) use.  Needs no device. */
int fk20_twiddle_selfcheck(unsigned *total_ops, unsigned *wnaf_width);
/* Host-only copy of one program the transform kernels run: which = 0 for w^idx, 1 for
 * w^-idx (idx < 512), 2 for 2^-idx (idx < 10), for wNAF width wnaf (3..5).  Writes at most
 * max_ops 32-bit ops (the encoding of ladder.hip.h) and the program's length to n_ops.  Needs
 * no device.  No reference counterpart (the reference multiplies by twiddles with
 * double-and-add, g1p_mul.cu:18). */
int fk20_twiddle_program(int which, unsigned idx, unsigned wnaf, uint32_t *ops, unsigned max_ops, unsigned *n_ops);

/* Measurement helper: blocks x 256 lanes each issue 8 x iters dependent-free 32 x 32 -> 64-bit
 * multiply-adds (v_mad_u64_u32, the instruction every Fp product is built from); out receives
 * blocks * 256 words that depend on every result.  Time it to obtain the device's
 * multiply-add issue rate.  No reference counterpart. */
int fk20_mad_probe(uint64_t *out, unsigned blocks, unsigned iters, fk20_stream_t stream);
/* Measurement helper behind the choice of limb layout: blocks x 256 lanes, lane l starting
 * from x = xs[l mod 1024] (12 words, < 2p), run two independent chains of iters Montgomery
 * multiplications by y (< 2p) and write their product (12 words per lane, < 2p) to out.
 * limbs = 29: the library's 14 x 29-bit product scanning (radix 2^406); limbs = 32: 12 x 32-bit
 * CIOS (radix 2^384).  No reference counterpart (the reference multiplies with inline PTX,
 * fp_mul.cuh:13). */
int fk20_fp_mul_probe(int limbs, uint32_t *out, const uint32_t *xs, const uint32_t *y, unsigned blocks,
                      unsigned iters, fk20_stream_t stream);

/* ------------------------------------------------------------------ helpers */

/* Canonical affine form of count points.  Replaces g1a_fromG1p_wrapper (g1.cuh:98; g1p.cu:168). */
int fk20_g1p_to_affine(fk20_g1a_t *out, const fk20_g1p_t *in, size_t count, fk20_stream_t stream);
/* 48-byte compressed serialisation of count points (out: count * 48 bytes), the format of
 * FK20_OUT_COMPRESSED.  No reference counterpart (the reference stores g1p_t, toByte.py:41). */
int fk20_g1p_compress(uint8_t *out, const fk20_g1p_t *in, size_t count, fk20_stream_t stream);

/* Element-wise arithmetic over count elements, the counterparts of the reference's test
 * wrappers (g1test_fft.cu:39 g1p_add_wrapper, :70 g1p_mul_wrapper; fptest/frtest kernels).
 * fp ops: 0 add, 1 sub, 2 mul, 3 sqr, 4 neg(a), 5 inv(a), 6 reduce(a)
 * fr ops: 0 add, 1 sub, 2 mul, 3 sqr, 4 neg(a), 5 inv(a), 6 reduce(a)
 * g1 ops: 0 add(p,q), 1 sub(p,q), 2 dbl(p), 3 mul(p,k), 4 neg(p), 5 add_affine(p, q as g1a),
 *         6 is_point(p) -> out x[0] = 0/1, 7 eq(p,q) -> out x[0] = 0/1 */
int fk20_fp_op(int op, uint64_t *out, const uint64_t *a, const uint64_t *b, size_t count,
               fk20_stream_t stream);
int fk20_fr_op(int op, fk20_fr_t *out, const fk20_fr_t *a, const fk20_fr_t *b, size_t count,
               fk20_stream_t stream);
int fk20_g1p_op(int op, fk20_g1p_t *out, const fk20_g1p_t *p, const void *q, const fk20_fr_t *k,
                size_t count, fk20_stream_t stream);

/* Device memory and stream plumbing for hosts without a GPU framework (ctypes, cgo). */
int fk20_device_count(int *count);
int fk20_set_device(int device);
int fk20_malloc(void **ptr, size_t bytes);
int fk20_free(void *ptr);
int fk20_memcpy_h2d(void *dst, const void *src, size_t bytes);
int fk20_memcpy_d2h(void *dst, const void *src, size_t bytes);
int fk20_memset(void *dst, int value, size_t bytes);
int fk20_memcpy_d2d(void *dst, const void *src, size_t bytes);
int fk20_synchronize(void);
int fk20_stream_create(fk20_stream_t *stream);
int fk20_stream_destroy(fk20_stream_t stream);
int fk20_stream_synchronize(fk20_stream_t stream);
typedef void *fk20_event_t; /* hipEvent_t */
int fk20_event_create(fk20_event_t *event);
int fk20_event_destroy(fk20_event_t event);
int fk20_event_record(fk20_event_t event, fk20_stream_t stream);
int fk20_event_elapsed(float *ms, fk20_event_t start, fk20_event_t stop); /* waits for stop */

#ifdef __cplusplus
}
#endif
#endif
