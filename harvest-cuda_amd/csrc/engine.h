// Host engine of the MI355X Harvest (counterpart of the reference's
// src/engine.h:1-35).  It derives WORLD's analysis parameters, calls the
// gfx950 backend through its C ABI (include/harvest_hip.h), and runs the
// sequential contour stages on the host in O(F).
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/harvest_hip.h"
#include "host_memory.h"

namespace harvest {

struct Result {
  std::vector<double> f0, times;
};

// Nonzero refined candidates in slot order, rows offsets[i]..offsets[i+1].
// With `pending` set, f0 and scores are still being downloaded into the
// workspace's pinned buffers: readers call hv_wait_candidates first.
struct CandidateRows {
  const int32_t* offsets;
  const double* f0;
  const double* scores;
  hv_workspace* pending = nullptr;
};

using Timings = std::map<std::string, double>;

struct Params {
  int bands, ratio, y_length, fft_size, frames, out_frames, capacity;
  double fs_work;
};

// WORLD's parameter derivation (harvest.cpp:1149-1181, 1219-1228).
Params derive_params(int64_t length, int fs, double f0_floor, double f0_ceil,
                     double frame_period);
// Validation of the public entry points (reference src/bindings.cpp:18-39);
// validate_params checks everything but the sample values, validate_samples
// only those (the backend checks them too, while staging host input).
void validate_params(int64_t length, int fs, double f0_floor, double f0_ceil,
                     double frame_period);
void validate_samples(const double* x, int64_t length);
void validate(const double* x, int64_t length, int fs, double f0_floor, double f0_ceil,
              double frame_period);

// Plain analysis of one waveform on the leased workspace.
Result analyze(hv_workspace* ws, const double* x, int64_t length, int fs, double f0_floor,
               double f0_ceil, double frame_period, int host_threads = 1,
               Timings* timings = nullptr);
// Batched analysis: one device launch sequence for all utterances, then the
// host contour of every utterance on a pool of host threads.  Given more
// workspaces on the same device, a large batch (see pipeline_groups) is cut
// into length buckets analysed in turn on the workspaces, so that the host
// staging and contours of some buckets overlap the device work of others.
// Results do not depend on the split.
std::vector<Result> analyze_batch(hv_workspace* ws, const std::vector<const double*>& xs,
                                  const std::vector<int64_t>& lengths, int fs,
                                  double f0_floor, double f0_ceil, double frame_period,
                                  int host_threads, Timings* timings = nullptr,
                                  const std::vector<hv_workspace*>& more = {});

// Batched analysis of waveforms already in device memory (back to back).
std::vector<Result> analyze_batch_device(hv_workspace* ws, const double* x_device,
                                         const std::vector<int64_t>& lengths, int fs,
                                         double f0_floor, double f0_ceil, double frame_period,
                                         int host_threads, Timings* timings = nullptr,
                                         const std::vector<hv_workspace*>& more = {});

// How a large batch is cut into length buckets (pipeline groups): a batch
// of at least 4 utterances and `split_frames` 1 ms frames is split into up
// to `max_groups` groups of at least `min_group_frames` frames; groups take
// shares of the frames in proportion to their weights (the first and the
// last group `first_weight` and `last_weight`, the others 1), so that a
// small first group starts the device early and a small last group shortens
// the host contour left after the device finishes.  Process-wide; the
// package applies its autotuned table (harvest_cuda/tuning.json).
struct BatchTuning {
  int max_groups = 6;
  int64_t min_group_frames = 50000;
  int64_t split_frames = 200000;
  double first_weight = 1.0;
  double last_weight = 1.0;
  // 1: a group's device work waits for the previous group's (the two
  // workspaces' kernels do not run concurrently; uploads, downloads and host
  // work still overlap the device).
  int serialize = 0;
  // Workspaces a pipelined batch keeps in flight (2-8): groups g, g + 1, ...,
  // g + depth - 1 are enqueued while group g's contours run.
  int depth = 2;
};
void set_batch_tuning(const BatchTuning& t);
BatchTuning batch_tuning();

// Number of pipeline groups for a batch under the current BatchTuning.
int pipeline_groups(const std::vector<int64_t>& frames);

// Contour stages, exposed for the bit-exactness tests against WORLD.
// FixF0Contour (harvest.cpp:1027-1044) from SearchF0Base's output `base`.
// Voiced sections are extended on up to `threads` threads.
void fix_contour(const CandidateRows& rows, int frames, const double* base, double* out,
                 int threads = 1);
// SmoothF0Contour (harvest.cpp:1079-1113), bit-exact, O(F + sections * settle).
// Sections are smoothed on up to `threads` threads.
void smooth_contour(const double* f0, int frames, double* out, int threads = 1);
// FixF0Contour split after FixStep3 (`step3`) for the differentiable form;
// `out` equals fix_contour's output.
void fix_contour_split(const CandidateRows& rows, int frames, const double* base, double* step3,
                       double* out, int threads = 1);
// FixStep4 (harvest.cpp:1000-1022) alone, and its transpose at `step3`.
void fix_step4_apply(const double* step3, int frames, double* out);
void fix_step4_adjoint(const double* step3, int frames, const double* g_out, double* g_in);
// Transpose of smooth_contour at contour f0 (the map is linear for a fixed
// voiced/unvoiced pattern).
void smooth_adjoint(const double* f0, int frames, const double* g_out, double* g_in);
// Resampling to the output frame period (harvest.cpp:1246-1251).
void resample(const double* basic, int frames, double frame_period, int out_frames,
              double* f0, double* times);
// The 1 ms frame each output frame of resample reads.
void resample_indices(int frames, double frame_period, int out_frames, int* index);

// Throws std::invalid_argument / std::runtime_error for a failed backend call.
void throw_status(hv_status status);

}  // namespace harvest
