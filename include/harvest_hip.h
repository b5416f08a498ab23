/*
 * harvest_hip.h - C ABI of the MI355X (gfx950) Harvest backend.
 *
 * The backend replaces the reference's C++ GPU boundary (src/engine.h:20-34,
 * implemented in src/cuda_backend.cu).  Every entry point:
 *   - takes plain pointers and sizes (no C++ or torch types);
 *   - catches everything internally and returns an hv_status; the message of
 *     the last failure on the calling thread is available from hv_last_error();
 *   - saves and restores the caller's current HIP device (the HIP counterpart
 *     of CudaContextGuard, src/cuda_device.h:92-108);
 *   - borrows host pointers only for the duration of the call.  Device memory
 *     belongs to the workspace and is reused (grow-only) across calls.
 * A workspace must be used by one thread at a time; the host-side
 * GpuWorkspaceCache lease (workspace_cache.h) guarantees this.
 */
#ifndef HARVEST_HIP_H_
#define HARVEST_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hv_workspace hv_workspace;

typedef enum hv_status {
  HV_OK = 0,
  HV_EINVAL = 1,       /* invalid argument: maps to ValueError             */
  HV_ENODEV = 2,       /* no usable GPU: maps to RuntimeError              */
  HV_EUNSUPPORTED = 3, /* GPU present but not gfx950                       */
  HV_ENOMEM = 4,       /* device or pinned host allocation failed          */
  HV_EDEVICE = 5,      /* HIP runtime error                                */
  HV_EINTERNAL = 6     /* anything else                                    */
} hv_status;

/* Timings (milliseconds, hipEvents on the workspace stream) written by the
 * analysis entry points when `timings_ms` is not NULL.  Stage spans:
 *   [0] preprocess     upload, decimation, mean removal
 *   [1] filter_events  two-bin terms, filterbank, crossings, raw and official
 *                      candidates
 *   [2] refine         refinement, pruning, compaction, download
 * Kernel sums inside those spans:
 *   [3] decimation scan kernels      [4] filterbank + crossing kernel
 *   [5] tile scan + interpolation    [6] refinement kernels
 *   [7] prune, best, scan, compaction
 *   [8] bin-table DFT kernel alone, its bins and the slots' FixF0 (part
 *       of [6]; 0 when the per-frame refinement ran) */
#define HV_TIMING_SLOTS 9

/* Work counters of the last analysis call (hv_last_stats):
 *   [0] refinement window samples computed (distinct (frame, h) windows)
 *   [1] refinement bin-samples: one term of the main and of the derivative
 *       spectrum DFT at one bin (distinct (frame, h, k) bins)
 *   [2] refinement bins computed
 *   [3] (frame, h) windows computed by the bin-table refinement (0 when the
 *       per-frame refinement ran) */
#define HV_STAT_SLOTS 4

/* Tuning of a workspace.  No reference counterpart: the package's
 * autotuner (harvest_cuda/autotune.py) picks the values on the GPU and
 * applies its committed table to every workspace it creates.
 *   fft_min_half     bands whose filter half length is at least this, when
 *                    the longest band fits a 4096-point block, run the
 *                    overlap-save FFT filterbank; shorter ones the direct FIR
 *                    (default 64; a value above 1000 makes every band direct);
 *   graphs           1: each analysis' device launch sequence (decimation
 *                    through the download of offsets and best) is captured
 *                    into a HIP graph once per plan shape and buffer set and
 *                    replayed after that; analyses that collect timings run
 *                    launch by launch (default 0);
 *   slab_budget_mib  device memory of the crossing slabs of one band group
 *                    (default 4096; at least 64);
 *   refine_table     1: analyses whose 1 ms frames start at whole samples
 *                    (working rate a multiple of 1000 Hz) refine candidates
 *                    through per-half-window bin tables, one lane per (frame,
 *                    half window) pair; 0: one wavefront per frame for every
 *                    analysis (default 1);
 *   refine_groups    workgroups of the bin-table DFT kernel's persistent grid
 *                    (64 to 4096; 0: the default 1536).
 * Results do not depend on the tuning beyond the rounding of the two
 * filterbank forms (fft_min_half). */
typedef struct hv_tuning {
  int fft_min_half;
  int graphs;
  int slab_budget_mib;
  int refine_table;
  int refine_groups;
} hv_tuning;
hv_status hv_get_tuning(hv_workspace* ws, hv_tuning* out);
hv_status hv_set_tuning(hv_workspace* ws, const hv_tuning* in);
/* Graph launches that replayed a cached graph and graphs captured, since the
 * workspace was created.  No reference counterpart. */
hv_status hv_graph_stats(hv_workspace* ws, int64_t* replays, int64_t* captures);

/* Device and pinned buffers the workspaces of this process have allocated
 * so far, and how many of the device ones replaced (freed) a smaller buffer
 * of the same workspace.  Repeating a batch allocates none; a pipelined
 * batch on new workspaces frees none (hv_reserve_batches).  Either pointer
 * may be NULL.  No reference counterpart. */
void hv_allocation_stats(int64_t* allocations, int64_t* device_regrowths);

/* Message of the last failed call on this thread ("" if none). */
const char* hv_last_error(void);

/* 1 if `device` is a visible gfx950 GPU, else 0.  Never fails and leaves no
 * sticky error.  Replaces cuda_available (src/engine.h:28,
 * src/cuda_backend.cu:794-803). */
int hv_device_supported(int device);

/* Create / destroy the per-device workspace.  Replaces create_gpu_workspace
 * and delete_gpu_workspace (src/engine.h:21-22, src/cuda_backend.cu:804-817).
 * Errors: device < 0 -> HV_EINVAL ("nonnegative"); device out of range ->
 * HV_EINVAL ("out of range"); no GPU -> HV_ENODEV; not gfx950 ->
 * HV_EUNSUPPORTED (src/cuda_device.h:26-40). */
hv_status hv_workspace_create(int device, hv_workspace** out);
void hv_workspace_destroy(hv_workspace* ws);
int hv_workspace_device(const hv_workspace* ws);

/* WORLD's zero-phase Chebyshev decimation of x[length] by `ratio` (1..12)
 * into y[y_length], y_length = ceil(length / ratio), before mean removal.
 * Computed as an exact chunked linear-recurrence scan.  Replaces
 * gpu_decimate (src/engine.h:29, src/cuda_backend.cu:818-838). */
hv_status hv_decimate(hv_workspace* ws, const double* x, int64_t length,
                      int ratio, double* y, int64_t y_length);

/* hv_decimate for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc); it has no
 * counterpart in the reference.  x_device[length] and y_device[y_length] are
 * device memory on `device`; the kernels are queued on `stream` (a
 * hipStream_t, NULL for the null stream) and the call returns without
 * waiting: synchronise the stream before reading y_device.  No workspace is
 * used and nothing is allocated: the plan and the scan's intermediate arrays
 * live in one static 64 MiB device array per GPU, so inputs are limited to
 * about 7.6 million samples (longer ones return HV_EINVAL).  Calls that
 * overlap on one GPU share that array and race: serialise them.  Output is
 * bitwise equal to hv_decimate's. */
hv_status hv_decimate_device_noalloc(int device, const double* x_device, int64_t length,
                                     int ratio, double* y_device, int64_t y_length,
                                     void* stream);

/* Analysis of one decimated, mean-removed waveform y[y_length] at the working
 * rate fs_work: filterbank, crossings, raw/official candidates, refinement,
 * pruning and compaction.  Replaces gpu_analyze (src/engine.h:30-34,
 * src/cuda_backend.cu:1129-1147).  Outputs (host):
 *   offsets[frames + 1]  CSR row offsets of nonzero refined candidates,
 *   best[frames]         f0 of the best-scoring candidate per frame,
 *   *nnz                 offsets[frames].
 * The candidates themselves are fetched by hv_fetch_candidates (or read in
 * place through hv_result_view; offsets and best may then be NULL).  fft_size,
 * bands, boundaries and capacity follow WORLD's HarvestGeneralBody
 * (harvest.cpp:1149-1181); frames counts 1 ms frames. */
hv_status hv_analyze(hv_workspace* ws, const double* y, int64_t y_length,
                     int fft_size, double fs_work, const double* boundaries,
                     int bands, int frames, double f0_floor, double f0_ceil,
                     int capacity, int32_t* offsets, double* best,
                     int64_t* nnz, double* timings_ms);

/* Fused form used by the host engine: `count` raw waveforms at sample rate
 * fs go through decimation, mean removal and analysis without returning to
 * the host in between.  Utterance u has lengths[u] samples at xs[u]; its
 * frames[u] = int(1000 lengths[u] / fs) + 1 one-millisecond frames occupy rows
 * [frame_offset[u], frame_offset[u] + frames[u]) of the concatenated outputs,
 * with frame_offset the exclusive prefix sum of frames.  Outputs:
 *   offsets[total_frames + 1]  global CSR offsets (offsets[0] = 0),
 *   best[total_frames],
 *   *nnz = offsets[total_frames].
 * Every utterance's results are bitwise identical to a call with count = 1.
 * The waveforms are staged through pinned memory in slices (the copy of one
 * slice overlaps the upload of the one before); a sample that is infinite or
 * NaN returns HV_EINVAL ("x must contain only finite samples") before any
 * kernel is queued, as the reference's host validation does
 * (src/bindings.cpp:36-38).
 * This is the batched form (no reference counterpart) and, at count = 1,
 * the plain form; it fuses gpu_decimate, the host mean removal
 * (src/engine.cpp:278-281) and gpu_analyze. */
hv_status hv_analyze_batch(hv_workspace* ws, int count,
                           const double* const* xs, const int64_t* lengths,
                           int fs, double f0_floor, double f0_ceil,
                           int32_t* offsets, double* best, int64_t* nnz,
                           double* timings_ms);

/* hv_analyze_batch with the waveforms already resident in device memory:
 * x_device holds the `count` utterances back to back (lengths[u] samples
 * each) on the workspace's device, and must be complete before the call
 * (synchronise the producing stream first).  Outputs as hv_analyze_batch.
 * No reference counterpart: the reference always starts from host memory. */
hv_status hv_analyze_batch_device(hv_workspace* ws, int count, const double* x_device,
                                  const int64_t* lengths, int fs, double f0_floor,
                                  double f0_ceil, int32_t* offsets, double* best,
                                  int64_t* nnz, double* timings_ms);

/* hv_analyze_batch / hv_analyze_batch_device that return as soon as the
 * offsets, best and nnz are on the host, with the candidate download (f0,
 * then scores) still running on the workspace's stream.  Before reading the
 * candidates through hv_result_view, call hv_wait_candidates(ws, 0) for f0
 * and hv_wait_candidates(ws, 1) for the scores as well; hv_fetch_candidates
 * waits by itself.  No reference counterpart: the host engine runs contour
 * steps that need only `best` (FixStep1-2, harvest.cpp:710-762) and the
 * f0-only part of FixStep3 while the download completes. */
hv_status hv_analyze_batch_async(hv_workspace* ws, int count,
                                 const double* const* xs, const int64_t* lengths,
                                 int fs, double f0_floor, double f0_ceil,
                                 int32_t* offsets, double* best, int64_t* nnz,
                                 double* timings_ms);
hv_status hv_analyze_batch_device_async(hv_workspace* ws, int count, const double* x_device,
                                        const int64_t* lengths, int fs, double f0_floor,
                                        double f0_ceil, int32_t* offsets, double* best,
                                        int64_t* nnz, double* timings_ms);

/* Two-step form of hv_analyze_batch_async for pipelines that keep several
 * workspaces busy: _submit validates the batch and enqueues the whole
 * analysis on the workspace's stream (through the download of offsets, best
 * and the work counters) and returns at once; _collect waits for it, starts
 * the candidate download (as the _async entry points do) and returns
 * offsets, best and nnz.  Pass timings = 1 to _submit to collect timings_ms.
 * Until _collect, every other analysis on the workspace fails with
 * HV_EINVAL.  No reference counterpart. */
hv_status hv_analyze_batch_submit(hv_workspace* ws, int count, const double* const* xs,
                                  const int64_t* lengths, int fs, double f0_floor,
                                  double f0_ceil, int timings);
hv_status hv_analyze_batch_device_submit(hv_workspace* ws, int count, const double* x_device,
                                         const int64_t* lengths, int fs, double f0_floor,
                                         double f0_ceil, int timings);
hv_status hv_analyze_batch_collect(hv_workspace* ws, int32_t* offsets, double* best,
                                   int64_t* nnz, double* timings_ms);

/* Sizes every buffer of `ws` for the largest of `groups` batches at once:
 * group k holds counts[k] utterances, whose lengths follow those of group
 * k - 1 in `lengths`; host_input = 1 when they will come from host memory
 * (the _submit form), 0 for device input.  The buffers grow on the next
 * submission, once, to the largest size any group needs, so that submitting
 * the groups afterwards in any order neither allocates nor frees device or
 * pinned memory (a freed device buffer is cleared by the driver in the
 * background, which slows the device work that follows for a while).
 * Validates like hv_analyze_batch.  No reference counterpart. */
hv_status hv_reserve_batches(hv_workspace* ws, int groups, const int* counts,
                             const int64_t* lengths, int fs, double f0_floor, double f0_ceil,
                             int host_input);

/* The device work of the next analysis submitted on `ws` (after its input
 * uploads) starts only when everything enqueued on `before` so far has
 * finished: the host engine uses it to run the groups of a pipelined batch
 * on the device one after another (their uploads and the host work still
 * overlap), when the tuning asks for that.  No reference counterpart. */
hv_status hv_order_after(hv_workspace* ws, hv_workspace* before);

/* Wait until the last analysis' candidate f0 (part 0) or f0 and scores
 * (part 1) are in the pinned buffers hv_result_view exposes.  Returns at once
 * after the synchronous entry points.  No reference counterpart. */
hv_status hv_wait_candidates(hv_workspace* ws, int part);

/* Work counters (HV_STAT_SLOTS entries) of the last analysis on `ws`, for
 * roofline accounting.  No reference counterpart. */
hv_status hv_last_stats(hv_workspace* ws, int64_t* stats);

/* Copy the nnz candidates of the last hv_analyze / hv_analyze_batch call:
 * f0[nnz] and scores[nnz] in CSR order (frame order, then slot order). */
hv_status hv_fetch_candidates(hv_workspace* ws, double* f0, double* scores);

/* Zero-copy access to the last analysis' results in the workspace's pinned
 * host buffers: offsets[total_frames + 1], best[total_frames], f0[nnz],
 * scores[nnz].  Valid until the next call on `ws`.  The offsets/best outputs
 * of the analysis entry points may then be NULL.  No reference counterpart
 * (the reference copies into HostArrays, src/cuda_backend.cu:1117-1126). */
hv_status hv_result_view(hv_workspace* ws, const int32_t** offsets, const double** best,
                         const double** f0, const double** scores, int64_t* nnz);

/* Initial f0 of each candidate of the last analysis, in CSR order: the
 * official candidate (harvest.cpp:348-412) its slot was refined from, which
 * fixes the slot's window length, FFT size and harmonic bins.  No reference
 * counterpart: the input of the frozen-decision adjoint below. */
hv_status hv_fetch_initial(hv_workspace* ws, double* f0_initial);

/* Frozen-decision refinement and its adjoint, for the differentiable form
 * (the adjoint entry point proposed beside hv_analyze_batch; no reference
 * counterpart).  The waveforms are preprocessed exactly as in
 * hv_analyze_batch; then for each entry e (a 1 ms frame of the batch and the
 * initial f0 of a candidate there) the refined F0 of GetRefinedF0
 * (harvest.cpp:589-617) is evaluated with the window, bins and harmonic count
 * that the initial f0 fixes:
 *   refined[e]  (if refined != NULL) the refined F0;
 *   grad_x[ΣL]  (if grad_x != NULL, weight required) sum_e weight[e] times
 *               the gradient of refined[e] with respect to the waveforms,
 *               through refinement, mean removal and decimation.
 * frame[e] indexes the batch's concatenated 1 ms frames (utterance u's
 * frames start after those of utterances 0..u-1).  The gradient is summed
 * with atomics: its last bits may differ between runs. */
hv_status hv_refine_entries(hv_workspace* ws, int count, const double* const* xs,
                            const int64_t* lengths, int fs, double f0_floor, double f0_ceil,
                            int64_t n_entries, const int64_t* frame, const double* f0_initial,
                            const double* weight, double* refined, double* grad_x);

#ifdef __cplusplus
}
#endif
#endif /* HARVEST_HIP_H_ */
