/* fftahip.h -- C ABI of the MI355X (gfx950) transform library that replaces
 * FFTA.jl's complex fft/bfft path.
 *
 * FFTA.jl has no FFI: its boundary is the AbstractFFTs plan interface in
 * src/plan.jl.  Each entry point below names the reference function it
 * replaces (paths relative to the FFTA.jl source tree).  A Julia wrapper
 * reaches these through ccall; the Python package in ffta.jl_amd/fftahip
 * reaches them through ctypes.
 *
 * Conventions
 *  - Arrays are column-major, 1-based regions (as in Julia), interleaved
 *    complex: ComplexF32 = 2 x float (precision 32), ComplexF64 = 2 x double
 *    (precision 64).
 *  - direction: -1 = FFT_FORWARD (exp(-2 pi i jk/N)), +1 = FFT_BACKWARD
 *    (unnormalised inverse), as src/callgraph.jl:1.
 *  - Device pointers are plain pointers to device memory; `stream` is a
 *    hipStream_t passed as void* (NULL = the default stream).
 *  - Every call returns an fftahip_status; on failure fftahip_last_error()
 *    returns a thread-local message (the reference's exception text where
 *    the reference defines one).
 *  - A plan owns device scratch and tables; one execution of a given plan may
 *    be in flight at a time (FFTA plans share a workspace in the same way,
 *    src/callgraph.jl:40).  Plans are independent of each other.
 */
#ifndef FFTAHIP_H
#define FFTAHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  FFTAHIP_SUCCESS = 0,
  FFTAHIP_ERR_DIMENSION_MISMATCH = 1, /* Julia DimensionMismatch */
  FFTAHIP_ERR_ARGUMENT = 2,           /* Julia ArgumentError */
  FFTAHIP_ERR_DOMAIN = 3,             /* Julia DomainError */
  FFTAHIP_ERR_HIP = 4,                /* a HIP runtime call failed */
  FFTAHIP_ERR_ALLOC = 5,              /* device or host allocation failed */
  FFTAHIP_ERR_NOT_IMPLEMENTED = 6,
  FFTAHIP_ERR_CAPACITY = 7,           /* request exceeds a configured budget */
  FFTAHIP_ERR_NOT_READY = 8           /* serving: result not yet available */
} fftahip_status;

#define FFTAHIP_FORWARD (-1)
#define FFTAHIP_BACKWARD 1
#define FFTAHIP_C32 32
#define FFTAHIP_C64 64
#define FFTAHIP_DEFAULT_BLUESTEIN_CUTOFF 73

typedef struct fftahip_plan_s* fftahip_plan;

/* Library version string. */
const char* fftahip_version(void);
/* Text of a status code. */
const char* fftahip_status_string(fftahip_status s);
/* Message of the last failed call on this thread ("" if none). */
const char* fftahip_last_error(void);

/* plan_fft / plan_bfft (src/plan.jl:71-100, _plan_fft).
 * x has `ndims` dimensions of sizes dims[0..ndims-1] (column-major); region
 * lists the 1-based dimensions to transform (it is sorted, as _sort does).
 * One call graph is planned per region dimension (CallGraph{T}(size(x,r),
 * BLUESTEIN_CUTOFF), src/callgraph.jl:133-138); a length <= 0 fails with
 * DimensionMismatch("Array length must be strictly positive")
 * (src/callgraph.jl:83-85).  bluestein_cutoff mirrors the BLUESTEIN_CUTOFF
 * keyword (src/callgraph.jl:44); it selects FFTA's tree (reported by
 * fftahip_plan_describe) -- results agree either way within tolerance. */
fftahip_status fftahip_plan_create(fftahip_plan* plan, int precision, int direction, int ndims, const int64_t* dims,
                                   int nregion, const int64_t* region, int64_t bluestein_cutoff);

/* Base.size(p, i) (src/plan.jl:36-47): i < 1 -> DomainError, i > M -> 1. */
fftahip_status fftahip_plan_size(fftahip_plan plan, int i, int64_t* size_out);
/* Number of transformed dimensions M (the N of FFTAPlan_cx{T,N}). */
fftahip_status fftahip_plan_rank(fftahip_plan plan, int* rank_out);
fftahip_status fftahip_plan_precision(fftahip_plan plan, int* precision_out);
fftahip_status fftahip_plan_direction(fftahip_plan plan, int* direction_out);

/* Multiply every result of the plan by `scale`, fused into the pass that
 * writes the output (AbstractFFTs' ScaledPlan; ifft = bfft scaled by
 * 1/prod(size(x, region)), the normalisation AbstractFFTs applies on top of
 * FFTA's plan_bfft, src/plan.jl:74-75).  Default 1. */
fftahip_status fftahip_plan_set_scale(fftahip_plan plan, double scale);

/* The argument checks of LinearAlgebra.mul!(y, p, x) (src/plan.jl:157-164,
 * 170-182, 210-222, 260-272), with the plan's region as the caller holds it
 * (`region`, nregion entries; FFTA plans carry a mutable region).  Returns
 * FFTAHIP_ERR_DIMENSION_MISMATCH with the reference's message when the arrays
 * do not fit the plan. */
fftahip_status fftahip_plan_check(fftahip_plan plan, int nregion, const int64_t* region, int ndims_x,
                                  const int64_t* dims_x, int ndims_y, const int64_t* dims_y);

/* mul!(y, p, x) on device arrays of shape dims (src/plan.jl:157-167 for 1-D,
 * 170-207 for a 1-D plan on an N-D array, 210-257 and 260-325 for N-D).
 * Applies mul!'s argument checks first (as fftahip_plan_check with the
 * plan's region, x and y both of shape dims): a shape that does not fit the
 * plan returns FFTAHIP_ERR_DIMENSION_MISMATCH before anything is launched.
 * A pass over 2^31 or more lines is refused with FFTAHIP_ERR_CAPACITY.
 * Out of place; d_in may equal d_out.  Asynchronous on `stream`. */
fftahip_status fftahip_execute(fftahip_plan plan, int ndims, const int64_t* dims, const void* d_in, void* d_out,
                               void* stream);

/* The same on host arrays: copies in, transforms, copies out, synchronises.
 * This is the drop-in path for Julia Array arguments. */
fftahip_status fftahip_execute_host(fftahip_plan plan, int ndims, const int64_t* dims, const void* h_in,
                                    void* h_out);

/* fftahip_execute without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory -- code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  It
 * calls no hipMalloc/hipFree or other allocating routine; its scratch is a
 * fixed 64 MiB area inside the library on each GPU, held until the process
 * ends.  Limits: a transform needing more scratch than that fails with
 * FFTAHIP_ERR_CAPACITY; the array shape must already have a schedule on the
 * plan (the shape the plan was created with, or one run through
 * fftahip_execute before), else FFTAHIP_ERR_NOT_READY.  All calls on one GPU
 * share the area, which keeps the last call's contents: calls that overlap
 * on one GPU (concurrent streams or threads) race.  Same argument checks and
 * results as fftahip_execute. */
fftahip_status fftahip_execute_noalloc(fftahip_plan plan, int ndims, const int64_t* dims, const void* d_in,
                                       void* d_out, void* stream);

/* JSON description: FFTA's call-graph per region dimension (node type, size,
 * strides) and the device schedule (passes, radices, tiles). */
fftahip_status fftahip_plan_describe(fftahip_plan plan, char* buf, size_t cap, size_t* needed);

fftahip_status fftahip_plan_destroy(fftahip_plan plan);

/* ---------------------------------------------------------------- real transforms
 * plan_rfft(x, region) (src/plan.jl:102-125) for a real array x of shape
 * dims (direction -1; the result has size(x, region[1]) / 2 + 1 along the
 * first region dimension), or plan_brfft(x, len, region)
 * (src/plan.jl:127-150) for a complex array x of shape dims whose first
 * region dimension has len / 2 + 1 entries (direction +1; the result is real
 * with len entries there).  One or two region dimensions; more is
 * FFTAHIP_ERR_ARGUMENT "only supports 1D and 2D FFTs".  The multiplication
 * follows src/plan.jl:386-583 (split step for even lengths, complex
 * transform for odd ones; 2-D: d1 then d2 forward, d2 then d1 backward).
 * Real arrays are Float32/Float64 column-major; complex ones interleaved. */
typedef struct fftahip_rplan_s* fftahip_rplan;

fftahip_status fftahip_rplan_create(fftahip_rplan* plan, int precision, int direction, int ndims,
                                    const int64_t* dims, int64_t len, int nregion, const int64_t* region,
                                    int64_t bluestein_cutoff);
/* Shape of the result array (ndims_out entries written to dims_out if non-NULL). */
fftahip_status fftahip_rplan_output_dims(fftahip_rplan plan, int* ndims_out, int64_t* dims_out);
/* p * x on device arrays of the plan's shapes (asynchronous on stream). */
fftahip_status fftahip_rplan_execute(fftahip_rplan plan, const void* d_in, void* d_out, void* stream);
/* p * x on host arrays (staged through device memory; synchronous). */
fftahip_status fftahip_rplan_execute_host(fftahip_rplan plan, const void* h_in, void* h_out);
fftahip_status fftahip_rplan_destroy(fftahip_rplan plan);

/* ---------------------------------------------------------------- batches
 * Many independent 1-D transforms of uneven lengths (the batched use of
 * fft!(out, in, ...) per vector, src/algos.jl:5-34).  Transform i reads
 * lengths[i] elements at element offset offsets[i] of d_in and writes the same
 * range of d_out; directions[i] is -1 or +1.  offsets may be NULL (packed:
 * prefix sums of lengths). */
typedef struct fftahip_batch_s* fftahip_batch;

fftahip_status fftahip_batch_create(fftahip_batch* batch, int precision, int64_t count, const int64_t* lengths,
                                    const int64_t* offsets, const int* directions, int64_t bluestein_cutoff);
fftahip_status fftahip_batch_execute(fftahip_batch batch, const void* d_in, void* d_out, void* stream);
/* Builds the batch's launch DAG as a hipGraph on first use (kernel nodes
 * with the pass dependencies as edges, so independent launches overlap as in
 * fftahip_batch_execute) and replays it on later calls with the same
 * pointers. */
fftahip_status fftahip_batch_execute_graph(fftahip_batch batch, const void* d_in, void* d_out, void* stream);
fftahip_status fftahip_batch_describe(fftahip_batch batch, char* buf, size_t cap, size_t* needed);
fftahip_status fftahip_batch_destroy(fftahip_batch batch);

/* ---------------------------------------------------------------- serving
 * A request loop with continuous batching (no reference counterpart: FFTA
 * is a library call; this serves many fft/bfft calls concurrently).  One
 * dispatcher thread admits submitted requests in arrival order against a
 * device memory pool of pool_bytes (allocated once at creation; plan scratch,
 * descriptors and host-payload staging come from it), as soon as memory and
 * a stream slot are free, and runs them on `nstreams` non-blocking HIP
 * streams so that their work overlaps.  Requests of one plan (1-D, device
 * payloads, same precision, direction and length) pending together are
 * coalesced: up to 32 of them run as the lines of one batch, one set of
 * work-list launches.  Plans are cached per (precision, direction, shape)
 * and built outside the dispatcher's lock.  A request larger than the whole
 * pool fails with FFTAHIP_ERR_CAPACITY.  Payloads are device pointers, or
 * host pointers when host_buffers != 0 (staged through the pool).  Latency
 * runs from arrival (submission, or the scheduled time for replay) to
 * completion.  Env FFTAHIP_COALESCE=0 at creation turns coalescing off. */
typedef struct fftahip_server_s* fftahip_server;

typedef struct {
  int64_t completed, rejected, failed, pending;
  double p50_ms, p99_ms, mean_ms, max_ms;
  double seconds;          /* first arrival to last completion */
  double requests_per_s;   /* completed / seconds */
  double payload_gbps;     /* (input + output bytes) / seconds */
  int64_t pool_peak_bytes;
  int64_t plans;
  int64_t coalesced;       /* requests that ran inside coalesced groups */
  int64_t groups;          /* coalesced groups (one set of launches each) */
} fftahip_server_stats;

fftahip_status fftahip_server_create(fftahip_server* server, int64_t pool_bytes, int nstreams);
/* One request: the array of shape dims (column-major), region = all dims. */
fftahip_status fftahip_server_submit(fftahip_server server, int precision, int direction, int ndims,
                                     const int64_t* dims, const void* in, void* out, int host_buffers,
                                     uint64_t* ticket);
/* Submit `count` requests, request i at t_offset[i] seconds after the call
 * (a load generator for trace replay); dims holds 3 entries per request, so
 * ndims[i] is 1 to 3 (else FFTAHIP_ERR_ARGUMENT, nothing submitted). */
fftahip_status fftahip_server_replay(fftahip_server server, int64_t count, const double* t_offset,
                                     const int* precision, const int* direction, const int* ndims,
                                     const int64_t* dims, void* const* in, void* const* out, int host_buffers,
                                     uint64_t* tickets);
/* *done = 1 once the request finished (its status is then returned). */
fftahip_status fftahip_server_test(fftahip_server server, uint64_t ticket, int* done);
fftahip_status fftahip_server_wait(fftahip_server server, uint64_t ticket, double* latency_ms);
fftahip_status fftahip_server_drain(fftahip_server server);
fftahip_status fftahip_server_stats_get(fftahip_server server, fftahip_server_stats* stats);
fftahip_status fftahip_server_reset_stats(fftahip_server server);
fftahip_status fftahip_server_destroy(fftahip_server server);

/* ---------------------------------------------------------------- autotuner
 * No counterpart in the reference (FFTA picks radices by fixed rules,
 * src/callgraph.jl:76-126).  For each in-LDS line length in `lengths`
 * (13-smooth, up to the one-pass limit), time the candidate radix schedules
 * on this device with a batch of lines of that length and keep the fastest
 * for every plan created afterwards in this process; *changed_out counts the
 * lengths whose choice differs from the default rule.  The table can be saved
 * and loaded (text: "precision length radices..." per radix schedule,
 * "S precision length lds-bytes fewest-lines" per strided tile shape). */
fftahip_status fftahip_autotune_lines(int precision, int64_t count, const int64_t* lengths, int64_t* changed_out);
/* Tile shape of the strided passes (lines laid across memory: multi-pass
 * Stockham passes, strided region dimensions) for the transform lengths of
 * one array shape: times the plan of `dims` over `region` (forward) with
 * each candidate tile (LDS bytes per workgroup, fewest adjacent lines per
 * tile; the default rule is 40 KB tiles of >= 16 (CF32) / 8 (CF64) lines)
 * and keeps the fastest for those lengths in every later plan; a larger
 * tile must win by 3%.  *changed_out counts the lengths set to a non-default
 * shape.  Allocates two arrays of the shape on the device. */
fftahip_status fftahip_autotune_shape(int precision, int ndims, const int64_t* dims, int nregion,
                                      const int64_t* region, int64_t* changed_out);
fftahip_status fftahip_autotune_save(const char* path);
fftahip_status fftahip_autotune_load(const char* path, int64_t* entries_out);
fftahip_status fftahip_autotune_clear(void);
/* The radix schedule (stage order) plans use for an in-LDS line of `length`. */
fftahip_status fftahip_autotune_get(int precision, int64_t length, int* nrad, int* radices);
/* The strided tile shape plans use for a transform of `length` (0, 0: the
 * default rule). */
fftahip_status fftahip_autotune_get_tile(int precision, int64_t length, int64_t* lds_bytes, int64_t* min_lines);

/* ---------------------------------------------------------------- device helpers
 * Minimal memory helpers so that a host language without a GPU array type
 * (the Julia wrapper, plain C) can run the device entry points. */
fftahip_status fftahip_malloc(void** ptr, size_t bytes);
fftahip_status fftahip_free(void* ptr);
fftahip_status fftahip_memcpy_htod(void* dst, const void* src, size_t bytes);
fftahip_status fftahip_memcpy_dtoh(void* dst, const void* src, size_t bytes);
fftahip_status fftahip_synchronize(void);
/* Device properties the planner uses: LDS bytes per workgroup, CU count. */
fftahip_status fftahip_device_info(int64_t* lds_per_block, int64_t* cu_count, int64_t* warp_size);

/* ---------------------------------------------------------------- schedule inspection
 * Dry runs of the planner (no device needed): the launches that
 * fftahip_execute / fftahip_batch_execute would issue on a device with
 * `lds_per_block` bytes of LDS, as JSON with device pointers replaced by
 * [kind, offset] tokens (fftahip_debug_schedule plans with the given
 * BLUESTEIN_CUTOFF, as fftahip_plan_create does).  No reference counterpart;
 * used by the test-suite to check pass geometry on the host. */
fftahip_status fftahip_debug_schedule(int precision, int direction, int ndims, const int64_t* dims, int nregion,
                                      const int64_t* region, int64_t bluestein_cutoff, int64_t lds_per_block,
                                      char* buf, size_t cap, size_t* needed);
fftahip_status fftahip_debug_batch_schedule(int precision, int64_t count, const int64_t* lengths,
                                            const int* directions, int64_t lds_per_block, char* buf, size_t cap,
                                            size_t* needed);
/* Device read-back of the Bluestein chirp the plans use, b_n = exp(i pi p /
 * N), p = n^2 mod 2N folded into (-N, N] (src/algos.jl:353-359; the index
 * arithmetic is exact 64-bit integer work): entries n0 .. n0 + count - 1
 * for a length N < 2^31, as interleaved complex values of `precision` (the
 * FP64 value rounded once).  Needs a device. */
fftahip_status fftahip_debug_chirp(int precision, int64_t N, int64_t n0, int64_t count, void* host_out);
/* Device read-back of a register pass's FP64 chirp steps, d[t] = exp(i pi
 * ((t s mod 2N)^2 mod 2N) / N) for t < L (interleaved re, im).  Needs a
 * device. */
fftahip_status fftahip_debug_chirp_steps(int64_t N, int64_t s, int L, double* host_out);
/* Host evaluation of the register passes' 64-bit remainder m mod d (d <
 * 2^32, m < 2^62 with m / d < 2^31) through an FP64 quotient. */
fftahip_status fftahip_debug_mod_q31(const uint64_t* m, uint64_t d, uint64_t* r, int64_t count);
/* Host evaluation of the kernels' division by a runtime constant d. */
fftahip_status fftahip_debug_fastdiv(uint32_t d, const uint32_t* n, uint32_t* q, int64_t count);

#ifdef __cplusplus
}
#endif

#endif /* FFTAHIP_H */
