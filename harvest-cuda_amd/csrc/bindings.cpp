// Python module harvest_cuda._core (counterpart of the reference's nanobind
// module, src/bindings.cpp:1-105).  Same surface and argument semantics:
// x must be a 1-D C-contiguous float64 NumPy array (no conversion, read-only
// accepted), device is keyword-only, invalid arguments raise ValueError,
// missing GPUs RuntimeError, and the GIL is released for the whole analysis.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <thread>
#include <vector>

#include "adjoint.h"
#include "engine.h"
#include "workspace_cache.h"

namespace py = pybind11;
using namespace harvest;

namespace {

using Input = py::array_t<double, py::array::c_style>;
using Cache = WorkspaceCache<hv_workspace>;

// Cache keys: the device index, and the device index plus k kSecondSlot for
// the k-th extra workspace a pipelined batch uses on that device.
constexpr int kSecondSlot = 1 << 16;

// Leases of the extra workspaces of a pipelined batch (BatchTuning::depth).
struct ExtraLeases {
  std::vector<std::unique_ptr<WorkspaceCache<hv_workspace>::Lease>> leases;
  std::vector<hv_workspace*> ws;
};

// Backend tuning applied to every workspace the cache creates (and to the
// cached ones when it changes); set from the package's tuning table.
std::mutex g_tuning_mutex;
hv_tuning g_tuning{64, 0, 4096, 1, 0};

hv_tuning current_tuning() {
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  return g_tuning;
}

Cache& cache() {
  // Owned by the module for the life of the process; cleared explicitly.
  static Cache* instance = new Cache(
      [](int key) {
        hv_workspace* ws = nullptr;
        throw_status(hv_workspace_create(key % kSecondSlot, &ws));
        const hv_tuning t = current_tuning();
        const hv_status st = hv_set_tuning(ws, &t);
        if (st != HV_OK) {
          hv_workspace_destroy(ws);
          throw_status(st);
        }
        return ws;
      },
      [](hv_workspace* ws) { hv_workspace_destroy(ws); });
  return *instance;
}

ExtraLeases extra_workspaces(int device) {
  ExtraLeases e;
  const int depth = batch_tuning().depth;
  for (int k = 1; k < depth; ++k) {
    e.leases.push_back(std::make_unique<Cache::Lease>(cache(), device + k * kSecondSlot));
    e.ws.push_back(e.leases.back()->get());
  }
  return e;
}

const double* data_1d(const Input& x) {
  if (x.ndim() != 1) throw py::type_error("x must be a 1-D float64 array");
  return x.data();
}

py::array_t<double> to_numpy(std::vector<double>&& v) {
  auto* owner = new std::vector<double>(std::move(v));
  py::capsule capsule(owner, [](void* p) { delete static_cast<std::vector<double>*>(p); });
  return py::array_t<double>({(py::ssize_t)owner->size()}, {(py::ssize_t)sizeof(double)},
                             owner->data(), capsule);
}

int host_threads() {
  unsigned n = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(n ? n : 1u, 16u));
}

// Runs a call on host waveforms whose parameters are already validated.  The
// backend checks the sample values while it stages them for upload, so they
// are read once; if the call fails before that (no usable GPU, say), the
// samples are checked here so that a bad sample is still reported as such
// (ValueError), as when samples are checked before the device is touched.
template <class F>
auto samples_first(const std::vector<const double*>& ptrs, const std::vector<int64_t>& lengths,
                   F&& run) -> decltype(run()) {
  try {
    return run();
  } catch (const std::runtime_error&) {
    for (size_t u = 0; u < ptrs.size(); ++u) validate_samples(ptrs[u], lengths[u]);
    throw;
  }
}

}  // namespace

PYBIND11_MODULE(_core, m) {
  m.doc() = "MI355X (gfx950) Harvest F0 estimator backend";
  HostCache::instance().initialize();

  m.def("clear_cache", [] {
    py::gil_scoped_release release;
    cache().clear();
    HostCache::instance().clear();
  }, "Release cached host and GPU working memory across threads and devices. "
     "Workspaces in use are released when their calls finish.");

  m.def("cuda_available", [](int device) { return hv_device_supported(device) != 0; },
        py::arg("device") = 0,
        "Whether the selected visible GPU is a supported gfx950 device. "
        "Name kept for drop-in compatibility.");

  m.def("harvest", [](Input x, int fs, double floor, double ceil, double period, int device) {
    const double* p = data_1d(x);
    const int64_t n = (int64_t)x.size();
    validate_params(n, fs, floor, ceil, period);
    Result r = samples_first({p}, {n}, [&] {
      py::gil_scoped_release release;
      Cache::Lease ws(cache(), device);
      return analyze(ws.get(), p, n, fs, floor, ceil, period, host_threads());
    });
    return py::make_tuple(to_numpy(std::move(r.f0)), to_numpy(std::move(r.times)));
  }, py::arg("x").noconvert(), py::arg("fs"), py::arg("f0_floor") = 71.0,
     py::arg("f0_ceil") = 800.0, py::arg("frame_period") = 5.0, py::kw_only(),
     py::arg("device") = 0,
     "Estimate F0 on the GPU. x: contiguous 1-D float64 NumPy array; returns (f0, times).");

  m.def("harvest_batch", [](py::list xs, int fs, double floor, double ceil, double period,
                            int device, bool timings) {
    Timings t;
    std::vector<Input> keep;
    std::vector<const double*> ptrs;
    std::vector<int64_t> lengths;
    for (py::handle h : xs) {
      if (!Input::check_(h)) throw py::type_error("every utterance must be a 1-D C-contiguous float64 array");
      Input a = py::reinterpret_borrow<Input>(h);
      const double* p = data_1d(a);
      validate_params((int64_t)a.size(), fs, floor, ceil, period);
      keep.push_back(a);
      ptrs.push_back(p);
      lengths.push_back((int64_t)a.size());
    }
    if (ptrs.empty()) return py::object(py::list());
    std::vector<Result> rs = samples_first(ptrs, lengths, [&] {
      std::vector<Result> out;
      py::gil_scoped_release release;
      Cache::Lease ws(cache(), device);
      if (device >= kSecondSlot) throw std::invalid_argument("device index out of range");
      std::vector<int64_t> frames;
      for (int64_t n : lengths) frames.push_back(derive_params(n, fs, floor, ceil, period).frames);
      if (pipeline_groups(frames) > 1) {
        ExtraLeases more = extra_workspaces(device);
        out = analyze_batch(ws.get(), ptrs, lengths, fs, floor, ceil, period, host_threads(),
                            timings ? &t : nullptr, more.ws);
      } else {
        out = analyze_batch(ws.get(), ptrs, lengths, fs, floor, ceil, period, host_threads(),
                            timings ? &t : nullptr);
      }
      return out;
    });
    py::list out;
    for (auto& r : rs) out.append(py::make_tuple(to_numpy(std::move(r.f0)), to_numpy(std::move(r.times))));
    if (timings) return py::object(py::make_tuple(out, t));
    return py::object(out);
  }, py::arg("xs"), py::arg("fs"), py::arg("f0_floor") = 71.0, py::arg("f0_ceil") = 800.0,
     py::arg("frame_period") = 5.0, py::kw_only(), py::arg("device") = 0,
     py::arg("timings") = false,
     "Estimate F0 of many utterances in one batched GPU launch sequence.");

  m.def("harvest_device", [](uintptr_t x_device, std::vector<int64_t> lengths, int fs,
                             double floor, double ceil, double period, int device,
                             bool timings) {
    // Waveforms already resident in HBM (e.g. a torch float64 tensor's
    // data_ptr()); validation of sample values is the caller's business.
    // Parameter checks only; the samples stay on the device.
    for (int64_t n : lengths) validate_params(n, fs, floor, ceil, period);
    std::vector<Result> rs;
    Timings t;
    {
      py::gil_scoped_release release;
      Cache::Lease ws(cache(), device);
      if (device >= kSecondSlot) throw std::invalid_argument("device index out of range");
      std::vector<int64_t> frames;
      for (int64_t n : lengths) frames.push_back(derive_params(n, fs, floor, ceil, period).frames);
      const double* xd = reinterpret_cast<const double*>(x_device);
      if (pipeline_groups(frames) > 1) {
        ExtraLeases more = extra_workspaces(device);
        rs = analyze_batch_device(ws.get(), xd, lengths, fs, floor, ceil, period, host_threads(),
                                  timings ? &t : nullptr, more.ws);
      } else {
        rs = analyze_batch_device(ws.get(), xd, lengths, fs, floor, ceil, period, host_threads(),
                                  timings ? &t : nullptr);
      }
    }
    py::list out;
    for (auto& r : rs) out.append(py::make_tuple(to_numpy(std::move(r.f0)), to_numpy(std::move(r.times))));
    if (timings) return py::object(py::make_tuple(out, t));
    return py::object(out);
  }, py::arg("x_device"), py::arg("lengths"), py::arg("fs"), py::arg("f0_floor") = 71.0,
     py::arg("f0_ceil") = 800.0, py::arg("frame_period") = 5.0, py::kw_only(),
     py::arg("device") = 0, py::arg("timings") = false,
     "Batched analysis of waveforms already in device memory, stored back to back.");

  m.def("_profile", [](Input x, int fs, double floor, double ceil, double period, bool gpu,
                       int device) {
    if (!gpu) throw std::invalid_argument("this build profiles the GPU path only (gpu=True)");
    const double* p = data_1d(x);
    const int64_t n = (int64_t)x.size();
    validate_params(n, fs, floor, ceil, period);
    Timings t;
    Result r = samples_first({p}, {n}, [&] {
      py::gil_scoped_release release;
      Cache::Lease ws(cache(), device);
      return analyze(ws.get(), p, n, fs, floor, ceil, period, host_threads(), &t);
    });
    return py::make_tuple(to_numpy(std::move(r.f0)), to_numpy(std::move(r.times)), t);
  }, py::arg("x").noconvert(), py::arg("fs"), py::arg("f0_floor") = 71.0,
     py::arg("f0_ceil") = 800.0, py::arg("frame_period") = 5.0, py::arg("gpu") = true,
     py::kw_only(), py::arg("device") = 0);

  // ---- stage access for parity tests (device results, no host contour) ----
  m.def("_decimate", [](Input x, int fs, int device) {
    const double* p = data_1d(x);
    const int64_t n = (int64_t)x.size();
    validate(p, n, fs, 71.0, 800.0 < fs / 2.0 ? 800.0 : fs / 4.0, 5.0);
    const Params prm = derive_params(n, fs, 71.0, 800.0, 5.0);
    std::vector<double> y(prm.y_length);
    {
      py::gil_scoped_release release;
      Cache::Lease ws(cache(), device);
      throw_status(hv_decimate(ws.get(), p, n, prm.ratio, y.data(), prm.y_length));
    }
    return to_numpy(std::move(y));
  }, py::arg("x").noconvert(), py::arg("fs"), py::kw_only(), py::arg("device") = 0,
     "Device decimation before mean removal (WORLD decimate()).");

  m.def("_candidates", [](py::list xs, int fs, double floor, double ceil, int device) {
    std::vector<Input> keep;
    std::vector<const double*> ptrs;
    std::vector<int64_t> lengths;
    int64_t frames = 0;
    for (py::handle h : xs) {
      Input a = py::reinterpret_borrow<Input>(h);
      ptrs.push_back(data_1d(a));
      lengths.push_back((int64_t)a.size());
      validate(ptrs.back(), lengths.back(), fs, floor, ceil, 5.0);
      frames += derive_params(lengths.back(), fs, floor, ceil, 5.0).frames;
      keep.push_back(a);
    }
    std::vector<int32_t> offsets(frames + 1);
    std::vector<double> best(frames), f0, scores;
    int64_t nnz = 0;
    {
      py::gil_scoped_release release;
      Cache::Lease ws(cache(), device);
      throw_status(hv_analyze_batch(ws.get(), (int)ptrs.size(), ptrs.data(), lengths.data(), fs,
                                    floor, ceil, offsets.data(), best.data(), &nnz, nullptr));
      f0.resize(std::max<int64_t>(nnz, 1));
      scores.resize(std::max<int64_t>(nnz, 1));
      throw_status(hv_fetch_candidates(ws.get(), f0.data(), scores.data()));
      f0.resize(nnz);
      scores.resize(nnz);
    }
    py::array_t<int32_t> off(offsets.size());
    std::copy(offsets.begin(), offsets.end(), off.mutable_data());
    return py::make_tuple(off, to_numpy(std::move(f0)), to_numpy(std::move(scores)),
                          to_numpy(std::move(best)));
  }, py::arg("xs"), py::arg("fs"), py::arg("f0_floor") = 71.0, py::arg("f0_ceil") = 800.0,
     py::kw_only(), py::arg("device") = 0,
     "Device CSR of pruned refined candidates and the best f0 per 1 ms frame.");

  // ---- differentiable form (frozen decisions) ----
  py::class_<DiffRecord>(m, "_DiffRecord",
                         "Decisions of a recorded forward pass (frozen-decision mode).")
      .def_property_readonly("entries", [](const DiffRecord& r) { return r.entry_frame.size(); })
      .def_property_readonly("lengths", [](const DiffRecord& r) { return r.lengths; })
      .def_property_readonly("fs", [](const DiffRecord& r) { return r.fs; })
      .def_property_readonly("out_frames", [](const DiffRecord& r) {
        std::vector<int> v;
        for (const auto& u : r.utts) v.push_back(u.out_frames);
        return v;
      })
      .def_property_readonly("frame_period", [](const DiffRecord& r) { return r.frame_period; })
      .def_property_readonly("f0_floor", [](const DiffRecord& r) { return r.f0_floor; })
      .def_property_readonly("f0_ceil", [](const DiffRecord& r) { return r.f0_ceil; })
      .def_property_readonly("entry_frame", [](const DiffRecord& r) { return r.entry_frame; })
      .def_property_readonly("entry_init", [](const DiffRecord& r) { return r.entry_init; })
      .def_property_readonly("sources", [](const DiffRecord& r) {
        py::list out;
        for (const auto& u : r.utts) out.append(py::cast(u.source));
        return out;
      })
      .def_property_readonly("step3", [](const DiffRecord& r) {
        py::list out;
        for (const auto& u : r.utts) out.append(py::cast(u.step3));
        return out;
      })
      .def_property_readonly("frame_base", [](const DiffRecord& r) {
        std::vector<int64_t> v;
        for (const auto& u : r.utts) v.push_back(u.frame_base);
        return v;
      });

  auto collect = [](py::list xs, std::vector<Input>& keep, std::vector<const double*>& ptrs,
                    std::vector<int64_t>& lengths) {
    for (py::handle h : xs) {
      if (!Input::check_(h)) throw py::type_error("every utterance must be a 1-D C-contiguous float64 array");
      Input a = py::reinterpret_borrow<Input>(h);
      keep.push_back(a);
      ptrs.push_back(data_1d(a));
      lengths.push_back((int64_t)a.size());
    }
  };
  auto check_lengths = [](const DiffRecord& rec, const std::vector<int64_t>& lengths) {
    if (lengths != rec.lengths)
      throw std::invalid_argument("waveform lengths differ from the recorded forward pass");
  };

  m.def("_harvest_record", [collect](py::list xs, int fs, double floor, double ceil, double period,
                                     int device) {
    std::vector<Input> keep;
    std::vector<const double*> ptrs;
    std::vector<int64_t> lengths;
    collect(xs, keep, ptrs, lengths);
    if (ptrs.empty()) throw std::invalid_argument("no waveforms");
    for (size_t u = 0; u < ptrs.size(); ++u) validate(ptrs[u], lengths[u], fs, floor, ceil, period);
    auto rec = std::make_unique<DiffRecord>();
    std::vector<Result> rs;
    {
      py::gil_scoped_release release;
      Cache::Lease ws(cache(), device);
      rs = analyze_record(ws.get(), ptrs, lengths, fs, floor, ceil, period, host_threads(), *rec);
    }
    py::list out;
    for (auto& r : rs) out.append(py::make_tuple(to_numpy(std::move(r.f0)), to_numpy(std::move(r.times))));
    return py::make_tuple(out, std::move(rec));
  }, py::arg("xs"), py::arg("fs"), py::arg("f0_floor") = 71.0, py::arg("f0_ceil") = 800.0,
     py::arg("frame_period") = 5.0, py::kw_only(), py::arg("device") = 0,
     "Batched forward that also records its decisions: ([(f0, t)], record).");

  m.def("_harvest_frozen", [collect, check_lengths](const DiffRecord& rec, py::list xs, int device) {
    std::vector<Input> keep;
    std::vector<const double*> ptrs;
    std::vector<int64_t> lengths;
    collect(xs, keep, ptrs, lengths);
    check_lengths(rec, lengths);
    for (size_t u = 0; u < ptrs.size(); ++u)
      validate(ptrs[u], lengths[u], rec.fs, rec.f0_floor, rec.f0_ceil, rec.frame_period);
    std::vector<std::vector<double>> f0s;
    {
      py::gil_scoped_release release;
      Cache::Lease ws(cache(), device);
      f0s = frozen_forward(ws.get(), ptrs, rec);
    }
    py::list out;
    for (auto& f : f0s) out.append(to_numpy(std::move(f)));
    return out;
  }, py::arg("record"), py::arg("xs"), py::kw_only(), py::arg("device") = 0,
     "Output F0 of (perturbed) waveforms with the record's decisions held.");

  m.def("_harvest_adjoint", [collect, check_lengths](const DiffRecord& rec, py::list xs,
                                                     py::list g_f0, int device) {
    std::vector<Input> keep, gkeep;
    std::vector<const double*> ptrs, gptrs;
    std::vector<int64_t> lengths, glengths;
    collect(xs, keep, ptrs, lengths);
    collect(g_f0, gkeep, gptrs, glengths);
    check_lengths(rec, lengths);
    if (glengths.size() != rec.utts.size())
      throw std::invalid_argument("one output gradient per waveform");
    for (size_t u = 0; u < glengths.size(); ++u)
      if (glengths[u] != rec.utts[u].out_frames)
        throw std::invalid_argument("output gradient length differs from the output frames");
    for (size_t u = 0; u < ptrs.size(); ++u)
      validate(ptrs[u], lengths[u], rec.fs, rec.f0_floor, rec.f0_ceil, rec.frame_period);
    std::vector<std::vector<double>> gx;
    {
      py::gil_scoped_release release;
      Cache::Lease ws(cache(), device);
      gx = adjoint(ws.get(), ptrs, rec, gptrs);
    }
    py::list out;
    for (auto& g : gx) out.append(to_numpy(std::move(g)));
    return out;
  }, py::arg("record"), py::arg("xs"), py::arg("g_f0"), py::kw_only(), py::arg("device") = 0,
     "Gradient of sum(g_f0 * f0) with respect to the waveforms, decisions held.");

  // ---- host contour adjoints (CPU only) for tests ----
  m.def("_smooth_adjoint", [](Input f0, Input g) {
    if (g.size() != f0.size()) throw std::invalid_argument("length mismatch");
    std::vector<double> out(f0.size());
    smooth_adjoint(f0.data(), (int)f0.size(), g.data(), out.data());
    return to_numpy(std::move(out));
  });
  m.def("_fix_step4", [](Input step3) {
    std::vector<double> out(step3.size());
    fix_step4_apply(step3.data(), (int)step3.size(), out.data());
    return to_numpy(std::move(out));
  });
  m.def("_fix_step4_adjoint", [](Input step3, Input g) {
    if (g.size() != step3.size()) throw std::invalid_argument("length mismatch");
    std::vector<double> out(step3.size());
    fix_step4_adjoint(step3.data(), (int)step3.size(), g.data(), out.data());
    return to_numpy(std::move(out));
  });

  // ---- host contour stages (CPU only) for bit-exactness tests ----
  m.def("_fix_contour", [](py::array_t<int32_t, py::array::c_style> offsets, Input f0,
                           Input scores, Input base, int threads) {
    const int frames = (int)base.size();
    if ((int)offsets.size() != frames + 1) throw std::invalid_argument("offsets must have frames + 1 entries");
    if (f0.size() != scores.size()) throw std::invalid_argument("f0 and scores differ in length");
    const int32_t* o = offsets.data();
    if (o[0] != 0 || o[frames] != (int64_t)f0.size())
      throw std::invalid_argument("offsets must start at 0 and end at the candidate count");
    for (int i = 0; i < frames; ++i)
      if (o[i + 1] < o[i]) throw std::invalid_argument("offsets must be non-decreasing");
    std::vector<double> out(frames);
    CandidateRows rows{o, f0.data(), scores.data()};
    {
      py::gil_scoped_release release;
      fix_contour(rows, frames, base.data(), out.data(), std::max(1, threads));
    }
    return to_numpy(std::move(out));
  }, py::arg("offsets"), py::arg("f0"), py::arg("scores"), py::arg("base"), py::arg("threads") = 1);
  m.def("_smooth", [](Input f0, int threads) {
    std::vector<double> out(f0.size());
    {
      py::gil_scoped_release release;
      smooth_contour(f0.data(), (int)f0.size(), out.data(), std::max(1, threads));
    }
    return to_numpy(std::move(out));
  }, py::arg("f0"), py::arg("threads") = 1);
  m.def("_params", [](int64_t length, int fs, double floor, double ceil, double period) {
    Params p = derive_params(length, fs, floor, ceil, period);
    py::dict d;
    d["bands"] = p.bands;
    d["ratio"] = p.ratio;
    d["y_length"] = p.y_length;
    d["fft_size"] = p.fft_size;
    d["frames"] = p.frames;
    d["out_frames"] = p.out_frames;
    d["capacity"] = p.capacity;
    d["fs_work"] = p.fs_work;
    return d;
  });
  m.def("_set_tuning", [](int fft_min_half, int graphs, int slab_budget_mib, int refine_table,
                          int refine_groups) {
    const hv_tuning t{fft_min_half, graphs, slab_budget_mib, refine_table, refine_groups};
    if (fft_min_half < 1 || (graphs != 0 && graphs != 1) || slab_budget_mib < 64 ||
        (refine_table != 0 && refine_table != 1) ||
        (refine_groups != 0 && (refine_groups < 64 || refine_groups > 4096)))
      throw std::invalid_argument("invalid tuning");
    {
      std::lock_guard<std::mutex> lock(g_tuning_mutex);
      g_tuning = t;
    }
    // Cached workspaces take it on their next lease by being dropped now
    // (those in use finish their call first).
    py::gil_scoped_release release;
    cache().clear();
  }, py::arg("fft_min_half"), py::arg("graphs"), py::arg("slab_budget_mib"),
     py::arg("refine_table") = 1, py::arg("refine_groups") = 0,
     "Backend tuning of every workspace (include/harvest_hip.h hv_tuning).");
  m.def("_get_tuning", [] {
    const hv_tuning t = current_tuning();
    py::dict d;
    d["fft_min_half"] = t.fft_min_half;
    d["graphs"] = t.graphs;
    d["slab_budget_mib"] = t.slab_budget_mib;
    d["refine_table"] = t.refine_table;
    d["refine_groups"] = t.refine_groups;
    return d;
  });
  m.def("_set_batch_tuning", [](int max_groups, int64_t min_group_frames, int64_t split_frames,
                                double first_weight, double last_weight, int serialize,
                                int depth) {
    BatchTuning t;
    t.max_groups = max_groups;
    t.min_group_frames = min_group_frames;
    t.split_frames = split_frames;
    t.first_weight = first_weight;
    t.last_weight = last_weight;
    t.serialize = serialize;
    t.depth = depth;
    set_batch_tuning(t);
  }, py::arg("max_groups"), py::arg("min_group_frames"), py::arg("split_frames"),
     py::arg("first_weight"), py::arg("last_weight"), py::arg("serialize") = 0,
     py::arg("depth") = 2,
     "How large batches are cut into length buckets (engine.h BatchTuning).");
  m.def("_get_batch_tuning", [] {
    const BatchTuning t = batch_tuning();
    py::dict d;
    d["max_groups"] = t.max_groups;
    d["min_group_frames"] = t.min_group_frames;
    d["split_frames"] = t.split_frames;
    d["first_weight"] = t.first_weight;
    d["last_weight"] = t.last_weight;
    d["serialize"] = t.serialize;
    d["depth"] = t.depth;
    return d;
  });
  m.def("_graph_stats", [](int device) {
    int64_t replays = 0, captures = 0, r = 0, c = 0;
    py::gil_scoped_release release;
    for (int key : {device, device + kSecondSlot, device + 2 * kSecondSlot,
                    device + 3 * kSecondSlot}) {
      Cache::Lease ws(cache(), key);
      throw_status(hv_graph_stats(ws.get(), &r, &c));
      replays += r;
      captures += c;
    }
    return std::make_pair(replays, captures);
  }, py::arg("device") = 0, "(graph replays, graph captures) of the device's workspaces.");
  m.def("_allocation_stats", [] {
    int64_t allocations = 0, regrowths = 0;
    hv_allocation_stats(&allocations, &regrowths);
    return std::make_pair(allocations, regrowths);
  }, "(buffers allocated, device buffers freed on growth) by the backend's workspaces so far.");
  m.def("_host_cache_allocated", [] { return HostCache::instance().allocated(); });
  m.def("_cached_workspaces", [] { return cache().size(); });
}
