// Host engine (engine.cpp) against the CPU stand-in backend
// (fake_backend.cpp): the plain call, a one-group batch and a batch
// pipelined over two workspaces return identical F0 for every utterance;
// several threads run pipelined batches and plain calls at once (sharing
// the host thread pool); a device failure in the middle of a pipelined
// batch surfaces as an exception and leaves both workspaces usable.  Built
// and run plainly and under ThreadSanitizer and AddressSanitizer by
// tests/test_host_native.py.
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <thread>
#include <vector>

#include "engine.h"

extern "C" {
hv_workspace* fake_workspace_create(int device);
void fake_workspace_destroy(hv_workspace* ws);
void fake_fail_next_submits(hv_workspace* ws, int n);
}

using namespace harvest;

namespace {
int failures = 0;
void require(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++failures;
  }
}

struct Workload {
  int fs = 16000;
  std::vector<std::vector<double>> xs;
  std::vector<const double*> ptrs;
  std::vector<int64_t> lengths;
  std::vector<double> flat;  // back to back, for the device-input form
};

Workload make(unsigned seed, int count) {
  Workload w;
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> secs(0.5, 30.0), val(-1.0, 1.0);
  for (int u = 0; u < count; ++u) {
    const int64_t n = (int64_t)(secs(rng) * w.fs);
    std::vector<double> x(n, 0.0);
    x[0] = val(rng);
    x[n / 2] = val(rng);
    x[n - 1] = val(rng);
    w.xs.push_back(std::move(x));
  }
  for (auto& x : w.xs) {
    w.ptrs.push_back(x.data());
    w.lengths.push_back((int64_t)x.size());
    w.flat.insert(w.flat.end(), x.begin(), x.end());
  }
  return w;
}

bool same(const std::vector<Result>& a, const std::vector<Result>& b) {
  if (a.size() != b.size()) return false;
  for (size_t u = 0; u < a.size(); ++u)
    if (a[u].f0 != b[u].f0 || a[u].times != b[u].times) return false;
  return true;
}

std::vector<Result> plain(hv_workspace* ws, const Workload& w, int threads) {
  std::vector<Result> out;
  for (size_t u = 0; u < w.xs.size(); ++u)
    out.push_back(analyze(ws, w.ptrs[u], w.lengths[u], w.fs, 71.0, 800.0, 5.0, threads));
  return out;
}
}  // namespace

int main() {
  const Workload w = make(7, 40);
  std::vector<int64_t> frames;
  for (int64_t n : w.lengths) frames.push_back(derive_params(n, w.fs, 71.0, 800.0, 5.0).frames);
  require(pipeline_groups(frames) > 1, "the workload is split into pipeline groups");

  hv_workspace* a = fake_workspace_create(0);
  hv_workspace* b = fake_workspace_create(0);
  const std::vector<Result> ref = plain(a, w, 8);
  require(ref[0].f0.size() > 100, "plain results");
  int voiced = 0;
  for (double f : ref[3].f0) voiced += f > 0;
  require(voiced > 10, "voiced frames survive the contour");

  require(same(analyze_batch(a, w.ptrs, w.lengths, w.fs, 71.0, 800.0, 5.0, 16), ref),
          "one-group batch equals plain calls");
  require(same(analyze_batch(a, w.ptrs, w.lengths, w.fs, 71.0, 800.0, 5.0, 16, nullptr, {b}), ref),
          "pipelined batch equals plain calls");
  {
    hv_workspace* c = fake_workspace_create(0);
    hv_workspace* d = fake_workspace_create(0);
    require(same(analyze_batch(a, w.ptrs, w.lengths, w.fs, 71.0, 800.0, 5.0, 16, nullptr,
                               {b, c, d}), ref),
            "four workspaces in flight");
    fake_workspace_destroy(c);
    fake_workspace_destroy(d);
  }
  Timings t;
  require(same(analyze_batch(a, w.ptrs, w.lengths, w.fs, 71.0, 800.0, 5.0, 16, &t, {b}), ref),
          "pipelined batch with timings equals plain calls");
  {
    // Other bucket shapes and serialised groups give the same results.
    const BatchTuning saved = batch_tuning();
    BatchTuning bt;
    bt.max_groups = 9;
    bt.min_group_frames = 1000;
    bt.first_weight = 0.3;
    bt.last_weight = 0.5;
    bt.serialize = 1;
    set_batch_tuning(bt);
    require(pipeline_groups(frames) == 9, "nine buckets");
    require(same(analyze_batch(a, w.ptrs, w.lengths, w.fs, 71.0, 800.0, 5.0, 16, nullptr, {b}), ref),
            "serialised weighted buckets equal plain calls");
    set_batch_tuning(saved);
    bool threw = false;
    try {
      bt.first_weight = 0.0;
      set_batch_tuning(bt);
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    require(threw && batch_tuning().max_groups == saved.max_groups, "invalid tuning rejected");
  }
  require(t.count("groups") && t["groups"] > 1, "timings report the groups");
  require(same(analyze_batch_device(a, w.flat.data(), w.lengths, w.fs, 71.0, 800.0, 5.0, 16,
                                    nullptr, {b}), ref),
          "pipelined back-to-back batch equals plain calls");

  // Concurrent callers on their own workspaces share the host pool.
  {
    std::vector<std::thread> callers;
    std::vector<int> ok(4, 0);
    for (int k = 0; k < 4; ++k)
      callers.emplace_back([&, k] {
        hv_workspace* x = fake_workspace_create(0);
        hv_workspace* y = fake_workspace_create(0);
        bool good = true;
        for (int rep = 0; rep < 2; ++rep) {
          if (k == 3)
            good &= same(plain(x, w, 4), ref);
          else
            good &= same(analyze_batch(x, w.ptrs, w.lengths, w.fs, 71.0, 800.0, 5.0, 8, nullptr, {y}),
                         ref);
        }
        ok[k] = good;
        fake_workspace_destroy(x);
        fake_workspace_destroy(y);
      });
    for (auto& c : callers) c.join();
    require(ok[0] && ok[1] && ok[2] && ok[3], "concurrent calls return the same results");
  }

  // A failure in the middle of a pipelined batch propagates; both
  // workspaces are collected and work afterwards.
  fake_fail_next_submits(a, 0);
  fake_fail_next_submits(b, 1);
  bool threw = false;
  try {
    analyze_batch(a, w.ptrs, w.lengths, w.fs, 71.0, 800.0, 5.0, 16, nullptr, {b});
  } catch (const std::runtime_error& e) {
    threw = std::strstr(e.what(), "injected device failure") != nullptr;
  }
  require(threw, "the device failure surfaces as runtime_error");
  require(same(analyze_batch(a, w.ptrs, w.lengths, w.fs, 71.0, 800.0, 5.0, 16, nullptr, {b}), ref),
          "both workspaces are usable after the failure");
  fake_workspace_destroy(a);
  fake_workspace_destroy(b);
  if (failures) return 1;
  std::puts("engine tests passed");
  return 0;
}
