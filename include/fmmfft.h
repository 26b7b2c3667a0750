/*
 * fmmfft.h -- C ABI of the MI355X-native FMM-FFT (libfmmfft_amd.so).
 *
 * Forward (or unnormalised inverse) 1-D complex DFT of length N = P*M by the
 * Edelman-McCorquodale-Toledo FMM-FFT: the signal is split into P polyphase
 * sections, the P-1 fractional-shift (cotangent-kernel) matvecs are applied
 * with a 1-D Chebyshev fast multipole method, the sections are combined by a
 * P-point DFT and P length-M FFTs finish the transform.  All work runs as
 * hand-written gfx950 HIP kernels on device-resident buffers.
 *
 * The entry points replace the reference's plan/execute pair:
 *   fmmfft_plan_create  <- mfti()      src/fmmfft.c:337-382 (+ mftii :261-332)
 *   fmmfft_execute      <- mft()       src/fmmfft.c:980-1081 (mftint :384-978
 *                                      and fftw_execute :1080)
 *   fmmfft_plan_sizes   <- getwlen2() / getvlen()  src/fmmfft.c:1110-1183
 * The reference's P MPI ranks are the P blocks of one device buffer: the
 * caller passes all N points at once, in natural order, and receives all N
 * outputs in natural order (block r of the output is what rank r held after
 * mft()).  There is no myid, no MPI communicator and no caller-carved
 * workspace: the plan owns its tables and its default workspace.
 *
 * Every function returns an fmmfft_status; none aborts.  Executes are
 * stream-ordered and never synchronise the host.  One plan may be used by one
 * stream at a time with its own workspace; concurrent executes of the same
 * plan must pass distinct workspaces to fmmfft_execute_ws().
 */
#ifndef FMMFFT_H
#define FMMFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fmmfft_plan_s *fmmfft_plan;

/* Storage / arithmetic type of a plan. */
enum fmmfft_dtype {
    FMMFFT_C64 = 0,  /* complex64 storage, FP32 contractions, FP64 setup,
                        section sums, combine and inter-pass twiddles    */
    FMMFFT_C128 = 1, /* complex128 end to end                           */
    FMMFFT_C64_SPLIT = 2 /* FMMFFT_C64 with the M2L flips as split products:
                            each FP64 operator carried as FP32 hi + lo and
                            both applied (twice the flip MFMAs)           */
};

enum fmmfft_status {
    FMMFFT_OK = 0,
    FMMFFT_EINVAL = 1,  /* invalid (N, P, terms, leaf), dtype, pointer or count */
    FMMFFT_ENOMEM = 2,  /* a device allocation failed                          */
    FMMFFT_EHIP = 3,    /* a HIP runtime call failed (see fmmfft_last_hip_error) */
    FMMFFT_ENOTSUP = 4, /* valid parameters outside what the kernels support    */
    FMMFFT_ETOOBIG = 5  /* workspace exceeds FMMFFT_STATIC_SCRATCH_BYTES
                           (fmmfft_execute_static only)                      */
};

/* Direction: +1 forward X[k] = sum x[n] e^{-2 pi i nk/N}; -1 unnormalised
 * inverse (conj(F(conj x))).  The reference only runs dir = +1
 * (src/fmmfft.c:1260); its dir < 0 branch conjugates before its local FFT
 * (src/fmmfft.c:976-977 vs :1080) and is not a correct inverse. */
#define FMMFFT_FORWARD 1
#define FMMFFT_INVERSE (-1)

/* Sizes of a would-be plan without creating it (replaces getwlen2 /
 * getvlen, src/fmmfft.c:1110-1183).  table_bytes: device operator tables;
 * work_bytes: device workspace for ONE signal.  Either pointer may be NULL.
 * Validity rules (the reference checks none of them, src/fmmfft.c:1200-1212):
 * FMMFFT_EINVAL unless N % P == 0, P a power of two with 2 <= P <= 64,
 * M = N/P divisible by leaf, s = M/leaf a power of two >= 8,
 * 2 <= terms <= 32, 2 <= leaf <= 128; FMMFFT_ENOTSUP (valid, but outside
 * what the kernels take) unless leaf % 4 == 0, leaf * P <= 8192 (complex64)
 * / 4096 (complex128), M a power of two and N <= 2^30 (the local FFTs run
 * in up to three passes of <= 1024 points; offsets into one signal are
 * 32-bit). */
int fmmfft_plan_sizes(int64_t n, int32_t p_blocks, int32_t terms, int32_t leaf,
                      int32_t dtype, size_t *table_bytes, size_t *work_bytes);

/* Build a plan on the current HIP device (replaces mfti, src/fmmfft.c:337):
 * n = N, p_blocks = P (the reference's MPI world size), terms = Chebyshev
 * order (the reference's T), leaf = points per leaf box (the reference's B;
 * tree depth = log2(N / (P * leaf))).  All operator tables are built on the
 * device in FP64 (plus FP32 copies for FMMFFT_C64).  max_batch sizes the
 * plan's default workspace (signals per fmmfft_execute_batched call);
 * 0 means 1. */
int fmmfft_plan_create(fmmfft_plan *out, int64_t n, int32_t p_blocks, int32_t terms,
                       int32_t leaf, int32_t dtype, int32_t max_batch);

int fmmfft_plan_destroy(fmmfft_plan plan);

/* One transform (replaces mft, src/fmmfft.c:980).  d_in, d_out: device
 * pointers to N complex values (complex64 or complex128 per the plan),
 * natural order; d_in == d_out is allowed.  d_in must be 16-byte aligned
 * (FMMFFT_EINVAL otherwise; device allocations are), d_out aligned to its
 * element.  stream: a hipStream_t (NULL = default stream). */
int fmmfft_execute(fmmfft_plan plan, const void *d_in, void *d_out, int32_t dir,
                   void *stream);

/* count independent signals, signal i at d_in + i*N (and d_out + i*N), one
 * launch sequence; count <= the plan's max_batch. */
int fmmfft_execute_batched(fmmfft_plan plan, int64_t count, const void *d_in,
                           void *d_out, int32_t dir, void *stream);

/* One transform that allocates nothing and uses no plan workspace (no
 * counterpart in the reference: for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The workspace is a static device array of
 * FMMFFT_STATIC_SCRATCH_BYTES, one per device, held until the process ends;
 * plans whose one-signal workspace (fmmfft_plan_sizes' work_bytes) exceeds it
 * are rejected with FMMFFT_ETOOBIG.  Every call on a device shares that
 * device's array: calls that overlap on one device (different streams or
 * threads without ordering between them) race.  Nothing in the array needs
 * clearing: each execute writes every word it reads before reading it.
 * Results equal fmmfft_execute's bit for bit. */
#define FMMFFT_STATIC_SCRATCH_BYTES (64u << 20)
int fmmfft_execute_static(fmmfft_plan plan, const void *d_in, void *d_out, int32_t dir, void *stream);

/* Bytes of caller workspace for count signals, and the execute that uses
 * it (d_work must be 256-byte aligned device memory of that size). */
size_t fmmfft_workspace_bytes(fmmfft_plan plan, int64_t count);
int fmmfft_execute_ws(fmmfft_plan plan, int64_t count, const void *d_in, void *d_out,
                      int32_t dir, void *d_work, size_t work_bytes, void *stream);

/* Plan parameters (any pointer may be NULL). */
int fmmfft_plan_get(fmmfft_plan plan, int64_t *n, int32_t *p_blocks, int32_t *terms,
                    int32_t *leaf, int32_t *dtype, int32_t *depth);

/* Copy a device operator table (always FP64) to host memory, for parity
 * checks.  which: 0 cheb [terms], 1 initch [leaf][terms], 2 level operators
 * [depth-2][8][terms][terms] (shftfn, shftfp, flip2n, flip2p, flip3n, flip3p,
 * shftln, shftlp), 3 topflp [terms][terms], 4 evalm [P-1][leaf][terms],
 * 5 evalmh [P/2][terms], 6 near kernel [P-1][4*leaf], 7 combine factors
 * [P] complex (re, im).  bytes must equal the table's size. */
int fmmfft_read_table(fmmfft_plan plan, int32_t which, void *host_dst, size_t bytes);

/* Copy an intermediate of the most recent execute that used the plan's
 * default workspace (signal 0) to host memory; synchronises the device.
 * which: 0 section sums [P] complex128 ((i/M) * sum, entry 0 = 0),
 * 1 multipole tree, 2 local tree (both levels 2..depth, node-major,
 * [node][terms][P-1][2] real in the plan's precision),
 * 3 section cot sums psi [P][M] (row 0 holds x_0) in the plan's complex
 * type -- only while M <= 1024: for larger M the two-pass local FFT reuses
 * the buffer for its intermediate.  bytes must equal the buffer's size. */
int fmmfft_read_stage(fmmfft_plan plan, int32_t which, void *host_dst, size_t bytes);

/* Per-stage timing with HIP events recorded on each execute's own stream,
 * for benchmarks (no reference counterpart; the reference times whole mft()
 * calls with omp_get_wtime, src/fmmfft.c:1259-1263).  Stages:
 * 0 S2M + section sums, 1 upward pass, 2 downward pass, 3 leaf evaluation
 * (L2T + near field), 4 combine, 5 local FFT pass 1, 6 the remaining local FFT passes (none
 * for M <= 1024, one for M <= 2^20, two above).
 * profile_begin creates events for up to max_calls executes; profile_end
 * synchronises, writes the summed milliseconds of each stage over the
 * recorded executes and their count, and stops recording. */
#define FMMFFT_NSTAGES 7
int fmmfft_profile_begin(fmmfft_plan plan, int32_t max_calls);
int fmmfft_profile_end(fmmfft_plan plan, double *stage_ms, int32_t nstages, int64_t *calls);

/* ---------------------------------------------------------------------------
 * Request server (no reference counterpart: the reference runs one transform
 * per MPI job, src/fmmfft.c:1192-1288).  A scheduler thread admits requests
 * of mixed (N, P, terms, leaf, dtype, dir) as they arrive, keeps plans and
 * workspaces cached on the device under an HBM budget (evicting idle shapes,
 * least recently used first; a request that does not fit waits, at the head
 * of the queue, until completions free memory), batches same-shape requests
 * (up to 2^23 points per batch) into one launch sequence replayed from a
 * captured hipGraph that reads and writes every request's own buffers in
 * place (through a device table of their addresses; no copies), and
 * overlaps batches of different shapes on separate streams, up to 16
 * batches at once: it launches the oldest request whose shape has a free
 * slot.  Requests carry caller-owned device buffers (input 16-byte aligned,
 * output aligned to one complex point; in place allowed) that must stay
 * valid until fmmfft_server_wait returns; every result equals what
 * fmmfft_execute returns for the same input.  bytes_in_use / peak_bytes
 * count every device allocation the server makes for its shapes (plan
 * tables, slot workspaces, address tables), all charged to the budget.
 * Latency is submit to completion on the device (a timing event after each
 * batch), not to when the scheduler notices it.
 * ------------------------------------------------------------------------- */
typedef struct fmmfft_server_s *fmmfft_server;

typedef struct fmmfft_server_stats {
    int64_t completed, batches, plans_created, plans_evicted, graphs_captured;
    double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> completion latency  */
    double span_s;                          /* first submit -> last completion */
    double nominal_gflop;                   /* sum of 5 N log2 N, completed   */
    size_t bytes_in_use, peak_bytes;        /* server-owned device memory    */
} fmmfft_server_stats;

int fmmfft_server_create(fmmfft_server *out, size_t hbm_budget_bytes, int32_t max_batch);
/* Queue one transform; returns at once with a ticket.  stream: the caller's
 * hipStream_t that produces d_in (NULL = default stream): an event recorded
 * on it now is waited for on the device before d_in is read.  EINVAL /
 * ENOTSUP for parameters fmmfft_plan_create would reject; a shape that
 * cannot fit the budget even alone completes with FMMFFT_ENOMEM. */
int fmmfft_server_submit(fmmfft_server srv, int64_t n, int32_t p_blocks, int32_t terms, int32_t leaf,
                         int32_t dtype, int32_t dir, const void *d_in, void *d_out, void *stream,
                         uint64_t *ticket);
/* Block until the ticket's transform has finished on the device; returns its
 * status and latency.  A ticket can be waited for once. */
int fmmfft_server_wait(fmmfft_server srv, uint64_t ticket, double *latency_ms);
int fmmfft_server_drain(fmmfft_server srv);
int fmmfft_server_stats_get(fmmfft_server srv, fmmfft_server_stats *stats);
int fmmfft_server_reset_stats(fmmfft_server srv);
int fmmfft_server_destroy(fmmfft_server srv);

const char *fmmfft_strerror(int status);
/* The last failing HIP error code seen by the library (0 if none). */
int fmmfft_last_hip_error(void);

#ifdef __cplusplus
}
#endif
#endif
