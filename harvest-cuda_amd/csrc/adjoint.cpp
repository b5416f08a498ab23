// Differentiable form: decision recording, frozen-decision forward and the
// hand-written reverse pass (host part).  Frozen decisions: which candidate
// each FixStep3 contour frame copies, FixStep4's bridged gaps, the voiced
// sections of smoothing and the resampling indices.  With those held the
// output is refined F0 values pushed through affine and linear maps:
//   f0_out = Resample . Smooth . FixStep4 . Copy (refined(entries)),
// and the reverse pass applies the transposes in the opposite order, then
// hands the per-entry weights to the device adjoint (hv_refine_entries).
#include "adjoint.h"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace harvest {

std::vector<Result> analyze_record(hv_workspace* ws, const std::vector<const double*>& xs,
                                   const std::vector<int64_t>& lengths, int fs, double f0_floor,
                                   double f0_ceil, double frame_period, int host_threads,
                                   DiffRecord& rec) {
  const int count = (int)xs.size();
  rec = DiffRecord{};
  rec.fs = fs;
  rec.f0_floor = f0_floor;
  rec.f0_ceil = f0_ceil;
  rec.frame_period = frame_period;
  rec.lengths = lengths;
  rec.utts.resize(count);
  int64_t base = 0;
  for (int u = 0; u < count; ++u) {
    const Params p = derive_params(lengths[u], fs, f0_floor, f0_ceil, frame_period);
    rec.utts[u].frames = p.frames;
    rec.utts[u].out_frames = p.out_frames;
    rec.utts[u].frame_base = base;
    base += p.frames;
  }
  int64_t nnz = 0;
  throw_status(hv_analyze_batch(ws, count, xs.data(), lengths.data(), fs, f0_floor, f0_ceil,
                                nullptr, nullptr, &nnz, nullptr));
  const int32_t* offsets = nullptr;
  const double *best = nullptr, *cf0 = nullptr, *csc = nullptr;
  throw_status(hv_result_view(ws, &offsets, &best, &cf0, &csc, &nnz));
  std::vector<double> init((size_t)std::max<int64_t>(nnz, 1));
  throw_status(hv_fetch_initial(ws, init.data()));
  std::vector<Result> results(count);
  std::vector<int> entry_of(nnz, -1);  // CSR index -> entry
  for (int u = 0; u < count; ++u) {
    DiffUtterance& d = rec.utts[u];
    const int F = d.frames;
    CandidateRows rows{offsets + d.frame_base, cf0, csc};
    d.step3.assign(F, 0.0);
    std::vector<double> contour(F), basic(F);
    fix_contour_split(rows, F, best + d.frame_base, d.step3.data(), contour.data(),
                      std::max(1, host_threads / count));
    smooth_contour(contour.data(), F, basic.data());
    results[u].f0.resize(d.out_frames);
    results[u].times.resize(d.out_frames);
    resample(basic.data(), F, frame_period, d.out_frames, results[u].f0.data(),
             results[u].times.data());
    // Every nonzero FixStep3 value is a candidate of its own frame
    // (SearchF0Base and SelectBestF0 copy candidates); equal values from
    // different slots share window and bins, so the first match serves.
    d.source.assign(F, -1);
    for (int i = 0; i < F; ++i) {
      if (d.step3[i] == 0.0) continue;
      int found = -1;
      for (int k = rows.offsets[i]; k < rows.offsets[i + 1]; ++k)
        if (cf0[k] == d.step3[i]) {
          found = k;
          break;
        }
      if (found < 0)
        throw std::runtime_error("contour value at frame " + std::to_string(i) +
                                 " matches no candidate of its frame");
      if (entry_of[found] < 0) {
        entry_of[found] = (int)rec.entry_frame.size();
        rec.entry_frame.push_back(d.frame_base + i);
        rec.entry_init.push_back(init[found]);
      }
      d.source[i] = entry_of[found];
    }
  }
  return results;
}

namespace {

void check_inputs(const std::vector<const double*>& xs, const DiffRecord& rec) {
  if (xs.size() != rec.utts.size())
    throw std::invalid_argument("expected " + std::to_string(rec.utts.size()) + " waveforms");
}

}  // namespace

std::vector<std::vector<double>> frozen_forward(hv_workspace* ws,
                                                const std::vector<const double*>& xs,
                                                const DiffRecord& rec) {
  check_inputs(xs, rec);
  const int64_t ne = (int64_t)rec.entry_frame.size();
  std::vector<double> value((size_t)std::max<int64_t>(ne, 1));
  throw_status(hv_refine_entries(ws, (int)xs.size(), xs.data(), rec.lengths.data(), rec.fs,
                                 rec.f0_floor, rec.f0_ceil, ne, rec.entry_frame.data(),
                                 rec.entry_init.data(), nullptr, value.data(), nullptr));
  std::vector<std::vector<double>> out(rec.utts.size());
  for (size_t u = 0; u < rec.utts.size(); ++u) {
    const DiffUtterance& d = rec.utts[u];
    std::vector<double> step3(d.frames, 0.0), contour(d.frames), basic(d.frames), times(d.out_frames);
    for (int i = 0; i < d.frames; ++i)
      if (d.source[i] >= 0) step3[i] = value[d.source[i]];
    fix_step4_apply(step3.data(), d.frames, contour.data());
    smooth_contour(contour.data(), d.frames, basic.data());
    out[u].resize(d.out_frames);
    resample(basic.data(), d.frames, rec.frame_period, d.out_frames, out[u].data(), times.data());
  }
  return out;
}

std::vector<std::vector<double>> adjoint(hv_workspace* ws, const std::vector<const double*>& xs,
                                         const DiffRecord& rec,
                                         const std::vector<const double*>& g_f0) {
  check_inputs(xs, rec);
  if (g_f0.size() != rec.utts.size()) throw std::invalid_argument("one output gradient per waveform");
  std::vector<double> weight(rec.entry_frame.size(), 0.0);
  for (size_t u = 0; u < rec.utts.size(); ++u) {
    const DiffUtterance& d = rec.utts[u];
    const int F = d.frames;
    // Resampling: output frame j reads 1 ms frame index[j].
    std::vector<int> index(d.out_frames);
    resample_indices(F, rec.frame_period, d.out_frames, index.data());
    std::vector<double> g_basic(F, 0.0), g_contour(F), g_step3(F), contour(F);
    for (int j = 0; j < d.out_frames; ++j) g_basic[index[j]] += g_f0[u][j];
    // The recorded voiced pattern fixes the smoothing sections and the gaps.
    fix_step4_apply(d.step3.data(), F, contour.data());
    smooth_adjoint(contour.data(), F, g_basic.data(), g_contour.data());
    fix_step4_adjoint(d.step3.data(), F, g_contour.data(), g_step3.data());
    for (int i = 0; i < F; ++i)
      if (d.source[i] >= 0) weight[d.source[i]] += g_step3[i];
  }
  int64_t total = 0;
  for (int64_t n : rec.lengths) total += n;
  std::vector<double> gx((size_t)std::max<int64_t>(total, 1));
  throw_status(hv_refine_entries(ws, (int)xs.size(), xs.data(), rec.lengths.data(), rec.fs,
                                 rec.f0_floor, rec.f0_ceil, (int64_t)weight.size(),
                                 rec.entry_frame.data(), rec.entry_init.data(), weight.data(),
                                 nullptr, gx.data()));
  std::vector<std::vector<double>> out(rec.utts.size());
  int64_t at = 0;
  for (size_t u = 0; u < rec.utts.size(); ++u) {
    out[u].assign(gx.begin() + at, gx.begin() + at + rec.lengths[u]);
    at += rec.lengths[u];
  }
  return out;
}

}  // namespace harvest
