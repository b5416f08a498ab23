// Differentiable form of the Harvest path: the forward pass records every
// discrete decision, after which the output F0 is a smooth function of the
// waveforms (frozen-decision mode).  The reverse pass is hand-written: host
// transposes of resampling, smoothing and FixStep4, then the device adjoint
// of refinement, mean removal and decimation (hv_refine_entries).  No
// reference counterpart: the reference is forward-only.
#ifndef HARVEST_ADJOINT_H_
#define HARVEST_ADJOINT_H_

#include <cstdint>
#include <vector>

#include "engine.h"

namespace harvest {

// Decisions of one utterance.
struct DiffUtterance {
  int frames = 0;       // 1 ms frames
  int out_frames = 0;   // frames at the output period
  int64_t frame_base = 0;
  std::vector<int> source;     // per 1 ms frame: entry its FixStep3 value copies, or -1
  std::vector<double> step3;   // FixStep3 output of the recorded forward
};

struct DiffRecord {
  int fs = 0;
  double f0_floor = 0, f0_ceil = 0, frame_period = 0;
  std::vector<int64_t> lengths;
  std::vector<DiffUtterance> utts;
  // Entries: candidates some contour value copies (global 1 ms frame and
  // the initial f0 that fixes the candidate's window and bins).
  std::vector<int64_t> entry_frame;
  std::vector<double> entry_init;
};

// Batched forward that also records the decisions.
std::vector<Result> analyze_record(hv_workspace* ws, const std::vector<const double*>& xs,
                                   const std::vector<int64_t>& lengths, int fs, double f0_floor,
                                   double f0_ceil, double frame_period, int host_threads,
                                   DiffRecord& record);

// Output F0 of (perturbed) waveforms with the record's decisions held.
std::vector<std::vector<double>> frozen_forward(hv_workspace* ws,
                                                const std::vector<const double*>& xs,
                                                const DiffRecord& record);

// Gradient of sum_u sum_j g_f0[u][j] f0[u][j] with respect to the waveforms,
// at waveforms xs (the record's or perturbed ones), decisions held.
std::vector<std::vector<double>> adjoint(hv_workspace* ws, const std::vector<const double*>& xs,
                                         const DiffRecord& record,
                                         const std::vector<const double*>& g_f0);

}  // namespace harvest

#endif  // HARVEST_ADJOINT_H_
