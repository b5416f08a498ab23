// A CPU stand-in for the gfx950 backend's C ABI (include/harvest_hip.h), for
// tests of the host engine (engine.cpp) on the CPU, under ThreadSanitizer
// and AddressSanitizer.  It implements the entry points the engine calls
// with the same contracts: batches of utterances, CSR results in buffers
// owned by the workspace and valid until its next call, the two-step
// submit/collect form, and a candidate "download" that completes on a
// helper thread after the call returns (f0 first, then scores), so that
// reading candidates before hv_wait_candidates is a race the sanitizer can
// see.  The candidates are synthetic but deterministic functions of each
// utterance's samples: voiced sections with a smooth track, near-duplicate
// and octave candidates, and sparse unvoiced candidates, which exercise
// every contour stage.  A workspace can be told to fail its next submit.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "harvest_hip.h"

namespace {
thread_local std::string g_error;

struct Batch {
  std::vector<const double*> xs;
  std::vector<int64_t> lengths;
  int fs = 0;
};
}  // namespace

struct hv_workspace {
  int device = 0;
  std::vector<int32_t> offsets;
  std::vector<double> best, f0, scores, f0_final, scores_final;
  int64_t nnz = -1;
  Batch pending;
  bool submitted = false;
  int fail_next_submit = 0;
  // Length lists of the groups of the last hv_reserve_batches call: every
  // later submission must be one of them.
  std::vector<std::vector<int64_t>> reserved;
  std::thread download;
  std::mutex mutex;
  std::condition_variable cv;
  int parts_ready = 2;  // 0: none, 1: f0, 2: f0 and scores

  void finish_download() {
    if (download.joinable()) download.join();
  }
  ~hv_workspace() { finish_download(); }
};

namespace {

hv_status fail(hv_status s, const char* m) {
  g_error = m;
  return s;
}

// Deterministic synthetic analysis of one utterance (1 ms frames).
void synthesize(const double* x, int64_t length, int fs, std::vector<int32_t>& off,
                std::vector<double>& best, std::vector<double>& f0, std::vector<double>& sc) {
  const int frames = (int)(1000.0 * (double)length / fs) + 1;
  const double phase = std::fabs(x[0]) * 13.0 + std::fabs(x[length / 2]) * 7.0;
  const int shift = (int)(std::fabs(x[length - 1]) * 1000.0) % 211;
  for (int i = 0; i < frames; ++i) {
    const bool voiced = ((i + shift) / 250) % 3 != 0 && i > 0 && i < frames - 1;
    const double track = 120.0 + 60.0 * std::sin(2.0 * 3.141592653589793 * i / 700.0 + phase);
    double b = 0.0;
    if (voiced) {
      const double vals[4] = {track, track * 1.02, track * 0.5, track * 2.0};
      const double scs[4] = {5.0 + (i % 3), 3.0, 2.0, 1.0};
      for (int k = 0; k < 4; ++k) {
        f0.push_back(vals[k]);
        sc.push_back(scs[k]);
      }
      b = track;
    } else if (i % 7 == 0) {
      f0.push_back(300.0 + (i % 11));
      sc.push_back(1.0);
      b = 300.0 + (i % 11);
    }
    best.push_back(b);
    off.push_back((int32_t)f0.size());
  }
}

// Runs the batch and starts the candidate download on a helper thread.
hv_status run(hv_workspace* ws, const Batch& b, int32_t* offsets, double* best, int64_t* nnz,
              double* timings, bool wait) {
  ws->finish_download();
  ws->offsets.assign(1, 0);
  ws->best.clear();
  ws->f0_final.clear();
  ws->scores_final.clear();
  for (size_t u = 0; u < b.xs.size(); ++u) {
    // Row ends are global positions in the batch's candidate arrays.
    synthesize(b.xs[u], b.lengths[u], b.fs, ws->offsets, ws->best, ws->f0_final,
               ws->scores_final);
  }
  ws->nnz = ws->offsets.back();
  ws->f0.assign(ws->nnz + 1, -1.0);
  ws->scores.assign(ws->nnz + 1, -1.0);
  {
    std::lock_guard<std::mutex> lk(ws->mutex);
    ws->parts_ready = 0;
  }
  ws->download = std::thread([ws] {
    std::this_thread::sleep_for(std::chrono::microseconds(200));
    std::copy(ws->f0_final.begin(), ws->f0_final.end(), ws->f0.begin());
    {
      std::lock_guard<std::mutex> lk(ws->mutex);
      ws->parts_ready = 1;
    }
    ws->cv.notify_all();
    std::this_thread::sleep_for(std::chrono::microseconds(200));
    std::copy(ws->scores_final.begin(), ws->scores_final.end(), ws->scores.begin());
    {
      std::lock_guard<std::mutex> lk(ws->mutex);
      ws->parts_ready = 2;
    }
    ws->cv.notify_all();
  });
  if (wait) ws->finish_download();
  if (offsets) std::memcpy(offsets, ws->offsets.data(), ws->offsets.size() * sizeof(int32_t));
  if (best) std::memcpy(best, ws->best.data(), ws->best.size() * sizeof(double));
  *nnz = ws->nnz;
  if (timings)
    for (int k = 0; k < HV_TIMING_SLOTS; ++k) timings[k] = 0.0;
  g_error.clear();
  return HV_OK;
}

// True when no reservation was made or (lengths, count) is one of its groups.
bool reserved_group(const hv_workspace* ws, int count, const int64_t* lengths) {
  if (ws->reserved.empty()) return true;
  for (const auto& g : ws->reserved)
    if ((int)g.size() == count && std::equal(g.begin(), g.end(), lengths)) return true;
  return false;
}

Batch batch_of(int count, const double* const* xs, const int64_t* lengths, int fs) {
  Batch b;
  b.xs.assign(xs, xs + count);
  b.lengths.assign(lengths, lengths + count);
  b.fs = fs;
  return b;
}

Batch device_batch(int count, const double* x, const int64_t* lengths, int fs) {
  Batch b;
  b.fs = fs;
  int64_t at = 0;
  for (int u = 0; u < count; ++u) {
    b.xs.push_back(x + at);
    b.lengths.push_back(lengths[u]);
    at += lengths[u];
  }
  return b;
}

}  // namespace

extern "C" {

const char* hv_last_error(void) { return g_error.c_str(); }

// Test hook: the next `n` submits on ws fail with HV_EDEVICE.
void fake_fail_next_submits(hv_workspace* ws, int n) { ws->fail_next_submit = n; }
hv_workspace* fake_workspace_create(int device) {
  auto* ws = new hv_workspace;
  ws->device = device;
  return ws;
}
void fake_workspace_destroy(hv_workspace* ws) { delete ws; }

hv_status hv_analyze_batch_async(hv_workspace* ws, int count, const double* const* xs,
                                 const int64_t* lengths, int fs, double, double,
                                 int32_t* offsets, double* best, int64_t* nnz,
                                 double* timings_ms) {
  if (!ws || ws->submitted) return fail(HV_EINVAL, "bad workspace state");
  return run(ws, batch_of(count, xs, lengths, fs), offsets, best, nnz, timings_ms, false);
}

hv_status hv_analyze_batch_device_async(hv_workspace* ws, int count, const double* x_device,
                                        const int64_t* lengths, int fs, double, double,
                                        int32_t* offsets, double* best, int64_t* nnz,
                                        double* timings_ms) {
  if (!ws || ws->submitted) return fail(HV_EINVAL, "bad workspace state");
  return run(ws, device_batch(count, x_device, lengths, fs), offsets, best, nnz, timings_ms, false);
}

hv_status hv_analyze_batch_submit(hv_workspace* ws, int count, const double* const* xs,
                                  const int64_t* lengths, int fs, double, double, int) {
  if (!ws || ws->submitted) return fail(HV_EINVAL, "the workspace has a submitted analysis");
  if (ws->fail_next_submit > 0) {
    --ws->fail_next_submit;
    return fail(HV_EDEVICE, "injected device failure");
  }
  if (!reserved_group(ws, count, lengths)) return fail(HV_EINVAL, "submitted group was not reserved");
  ws->pending = batch_of(count, xs, lengths, fs);
  ws->submitted = true;
  return HV_OK;
}

hv_status hv_analyze_batch_device_submit(hv_workspace* ws, int count, const double* x_device,
                                         const int64_t* lengths, int fs, double, double, int) {
  if (!ws || ws->submitted) return fail(HV_EINVAL, "the workspace has a submitted analysis");
  if (ws->fail_next_submit > 0) {
    --ws->fail_next_submit;
    return fail(HV_EDEVICE, "injected device failure");
  }
  if (!reserved_group(ws, count, lengths)) return fail(HV_EINVAL, "submitted group was not reserved");
  ws->pending = device_batch(count, x_device, lengths, fs);
  ws->submitted = true;
  return HV_OK;
}

hv_status hv_reserve_batches(hv_workspace* ws, int groups, const int* counts,
                             const int64_t* lengths, int, double, double, int) {
  if (!ws || ws->submitted) return fail(HV_EINVAL, "bad workspace state");
  if (groups < 1 || !counts || !lengths) return fail(HV_EINVAL, "bad groups");
  ws->reserved.clear();
  for (int g = 0; g < groups; ++g) {
    ws->reserved.emplace_back(lengths, lengths + counts[g]);
    lengths += counts[g];
  }
  return HV_OK;
}

hv_status hv_analyze_batch_collect(hv_workspace* ws, int32_t* offsets, double* best,
                                   int64_t* nnz, double* timings_ms) {
  if (!ws || !ws->submitted) return fail(HV_EINVAL, "no submitted analysis to collect");
  ws->submitted = false;
  return run(ws, ws->pending, offsets, best, nnz, timings_ms, false);
}

hv_status hv_wait_candidates(hv_workspace* ws, int part) {
  if (!ws || (part != 0 && part != 1)) return fail(HV_EINVAL, "bad part");
  std::unique_lock<std::mutex> lk(ws->mutex);
  ws->cv.wait(lk, [&] { return ws->parts_ready > part; });
  return HV_OK;
}

hv_status hv_result_view(hv_workspace* ws, const int32_t** offsets, const double** best,
                         const double** f0, const double** scores, int64_t* nnz) {
  if (!ws || ws->nnz < 0) return fail(HV_EINVAL, "no analysis result to view");
  *offsets = ws->offsets.data();
  *best = ws->best.data();
  *f0 = ws->f0.data();
  *scores = ws->scores.data();
  *nnz = ws->nnz;
  return HV_OK;
}

hv_status hv_order_after(hv_workspace* ws, hv_workspace* before) {
  if (!ws || !before || ws == before) return fail(HV_EINVAL, "bad workspaces");
  return HV_OK;  // the stand-in runs each analysis at collect, in call order
}

hv_status hv_last_stats(hv_workspace* ws, int64_t* stats) {
  if (!ws || !stats) return fail(HV_EINVAL, "null pointer");
  for (int k = 0; k < HV_STAT_SLOTS; ++k) stats[k] = 0;
  return HV_OK;
}

}  // extern "C"
