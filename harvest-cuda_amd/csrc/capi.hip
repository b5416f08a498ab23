// Workspace, planning and C ABI of the gfx950 Harvest backend
// (include/harvest_hip.h).  Replaces the reference's GpuWorkspace and its
// gpu_* entry points (src/cuda_backend.cu:92-135, 794-838, 1129-1147) with a
// batch-native device pipeline: one launch sequence serves any number of
// utterances, and a plain call is a batch of one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/harvest_hip.h"
#include "device_check.h"
#include "host_copy.h"
#include "kernels.h"

namespace hv {
namespace {

thread_local std::string g_error;

// The HIP runtime as device_check.h's runtime policy.
struct HipRuntime {
  using Error = hipError_t;
  static constexpr Error kSuccess = hipSuccess;
  static constexpr Error kOutOfMemory = hipErrorOutOfMemory;
  static constexpr Error kNoDevice = hipErrorNoDevice;
  static constexpr Error kInsufficientDriver = hipErrorInsufficientDriver;
  static Error device_count(int* n) { return hipGetDeviceCount(n); }
  static Error properties(int device, std::string* arch, std::string* name) {
    hipDeviceProp_t prop;
    const Error e = hipGetDeviceProperties(&prop, device);
    if (e == hipSuccess) {
      *arch = prop.gcnArchName;
      *name = prop.name;
    }
    return e;
  }
  static Error get_device(int* d) { return hipGetDevice(d); }
  static Error set_device(int d) { return hipSetDevice(d); }
  static const char* error_name(Error e) { return hipGetErrorName(e); }
  static const char* error_string(Error e) { return hipGetErrorString(e); }
  static void clear_error() { (void)hipGetLastError(); }
};
using Device = DeviceOps<HipRuntime>;
using DeviceGuard = Device::Guard;

void check(hipError_t e, const char* what) {
  // hipMalloc reports exhaustion as hipErrorMemoryAllocation as well.
  Device::check(e == hipErrorMemoryAllocation ? hipErrorOutOfMemory : e, what);
}

// Device and pinned buffers allocated by every workspace of the process,
// and the device buffers among them that replaced a smaller one (freed on
// growth); hv_allocation_stats.
std::atomic<long long> g_allocations{0}, g_device_regrowths{0};

// While set, ensure() records the size it is asked for instead of
// allocating (reserve_batches); a buffer whose recorded size exceeds its
// allocation grows to that size, once, at its next ensure().
thread_local bool g_recording = false;

struct Recording {
  Recording() { g_recording = true; }
  ~Recording() { g_recording = false; }
};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  size_t want = 0;  // largest size recorded by a reservation
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* ensure(size_t n) {
    size_t need = std::max<size_t>(n, 1) * sizeof(T);
    if (g_recording) {
      want = std::max(want, need);
      return static_cast<T*>(p);
    }
    if (need > bytes || want > bytes) {
      need = std::max(need, want);
      if (p) g_device_regrowths.fetch_add(1, std::memory_order_relaxed);
      release();
      check(hipMalloc(&p, need), "hipMalloc");
      bytes = need;
      g_allocations.fetch_add(1, std::memory_order_relaxed);
    }
    return static_cast<T*>(p);
  }
  template <class T> T* get() const { return static_cast<T*>(p); }
};

struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  size_t want = 0;  // largest size recorded by a reservation
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* ensure(size_t n) {
    size_t need = std::max<size_t>(n, 1) * sizeof(T);
    if (g_recording) {
      want = std::max(want, need);
      return static_cast<T*>(p);
    }
    if (need > bytes || want > bytes) {
      need = std::max(need, want);
      release();
      check(hipHostMalloc(&p, need, hipHostMallocDefault), "hipHostMalloc");
      bytes = need;
      g_allocations.fetch_add(1, std::memory_order_relaxed);
    }
    return static_cast<T*>(p);
  }
  template <class T> T* get() const { return static_cast<T*>(p); }
};

// WORLD FilterForDecimate coefficients a0, a1, a2, b0, b1 for r = 2..12
// (matlabfunctions.cpp:29-113).  The table must equal WORLD's bit for bit.
const double kDecimation[13][5] = {
    {0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0},
    {0.041156734567757189, -0.42599112459189636, 0.041037215479961225,
     0.16797464681802227, 0.50392394045406674},
    {0.95039378983237421, -0.67429146741526791, 0.15412211621346475,
     0.071221945171178636, 0.21366583551353591},
    {1.4499664446880227, -0.98943497080950582, 0.24578252340690215,
     0.036710750339322612, 0.11013225101796784},
    {1.7610939654280557, -1.2554914843859768, 0.3237186507788215,
     0.021334858522387423, 0.06400457556716227},
    {1.9715352749512141, -1.4686795689225347, 0.3893908434965701,
     0.013469181309343825, 0.040407543928031475},
    {2.1225239019534703, -1.6395144861046302, 0.44469707800587366,
     0.0090366882681608418, 0.027110064804482525},
    {2.2357462340187593, -1.7780899984041358, 0.49152555365968692,
     0.0063522763407111993, 0.019056829022133598},
    {2.3236003491759578, -1.8921545617463598, 0.53148928133729068,
     0.0046331164041389372, 0.013899349212416812},
    {2.3936475118069387, -1.9873904075111861, 0.5658879979027055,
     0.0034818622251927556, 0.010445586675578267},
    {2.450743295230728, -2.06794904601978, 0.59574774438332101,
     0.0026822508007163792, 0.0080467524021491377},
    {2.4981398605924205, -2.1368928194784025, 0.62187513816221485,
     0.0021097275904709001, 0.0063291827714127002},
};

constexpr size_t kSlabBudget = size_t(4) << 30;  // bytes of crossing slabs
constexpr int kTwiddleMax = 1 << 20;

int host_round(double x) { return x > 0 ? (int)(x + 0.5) : (int)(x - 0.5); }

// GetSuitableFFTSize (common.cpp:51-54), evaluated with the host libm as WORLD.
int suitable_fft_size(int sample) {
  return (int)std::pow(2.0, (int)(std::log((double)sample) / kLog2) + 1.0);
}

// IIR state transition A^kChunk in extended precision.
void chunk_power(const double* co, double* out) {
  long double a[9] = {co[0], co[1], co[2], 1, 0, 0, 0, 1, 0};
  long double r[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, t[9];
  for (int step = 0; step < kChunk; ++step) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        t[i * 3 + j] = a[i * 3 + 0] * r[0 * 3 + j] + a[i * 3 + 1] * r[1 * 3 + j] +
                       a[i * 3 + 2] * r[2 * 3 + j];
    std::memcpy(r, t, sizeof(r));
  }
  for (int k = 0; k < 9; ++k) out[k] = (double)r[k];
}

}  // namespace
}  // namespace hv

// Device memory of the allocation-free decimation entry point: one array per
// GPU, held by the loaded code object until the process ends.  Host code
// reaches each device's copy only through the address hipGetSymbolAddress
// returns with that device current (scratch_base); the array's name in host
// code is not a device address.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace hv {
namespace {

std::mutex g_scratch_mutex;
std::vector<void*> g_scratch_base;  // per device; sized once by the first call

// Address of this device's scratch array; the caller has made `device`
// current.  Looked up once per device.
void* scratch_base(int device, int devices) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_base.empty()) g_scratch_base.assign(std::max(devices, 1), nullptr);
  if (device >= (int)g_scratch_base.size()) invalid("device is out of range");
  if (!g_scratch_base[device]) {
    void* ptr = nullptr;
    check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) throw Failure(HV_EDEVICE, "scratch array has no device address");
    g_scratch_base[device] = ptr;
  }
  return g_scratch_base[device];
}

}  // namespace
}  // namespace hv

using namespace hv;

struct hv_workspace {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // Kernel timers: (slot, start, stop) event pairs recorded only when a call
  // asks for timings.
  std::vector<hipEvent_t> timer_events;
  std::vector<int> timer_slots;
  size_t timers_used = 0;
  bool timing = false;
  DevBuf x, pass, states, carries, carry_tmp, y, partial, mean, partial3, sums3, quirk;
  DevBuf slab, counts, offs, totals, raw, official, count;
  DevBuf rf0, rsc, pf0, psc, best, row_nnz, offsets, scan_tmp;
  // The plan (utterances and six offset tables) in one device buffer, uploaded
  // with a single copy, and typed views into it.
  DevBuf plan;
  const Utt* d_utt = nullptr;
  const long long *d_c_off = nullptr, *d_t_off = nullptr, *d_f_off = nullptr;
  const long long *d_p_off = nullptr, *d_y_off = nullptr, *d_x_off = nullptr;
  DevBuf cinit;                                        // initial f0 of each CSR entry
  DevBuf full, gy, gx, ent_frame, ent_f0, ent_w, ent_out;  // frozen-decision adjoint
  DevBuf coef, coef_off, half, bounds, twiddle, wintab, work;
  // Filterbank by FFT: (cos, sin)(2 pi t / 4096), the spectra of the FFT
  // bands [0, fft_bands) (cached with the filters) and of the tiles, and the
  // blocks' lead D (the longest FFT band's half length).
  DevBuf tw4096, band_spec, tile_spec;
  bool tw4096_ready = false;
  // Two-level phase table of the filterbank's correction (size phase_n).
  DevBuf phase;
  long long phase_n = 0;
  int fft_bands = 0, fft_D = 0;
  PinnedBuf h_x, h_meta, h_offsets, h_best, h_cand, h_work, h_ent, h_gx;
  // Filter-bank cache key (working rate, floor, ceil).
  double key_fs = -1.0, key_floor = -1.0, key_ceil = -1.0;
  int key_fft_min = -1;
  std::vector<double> boundaries;
  std::vector<int> halves;
  int max_half = 0;
  int twiddle_n = -1;
  int wintab_h = -1;  // largest half window the window table covers
  DevBuf winpair;     // (w, d) of integral-offset windows (k_window_pairs)
  int winpair_h = -1; // largest half window it covers
  // Bin tables of the refinement (refine_table.hip) for the key below, and
  // the per-analysis buffers of that path.
  DevBuf bt_info, bt_harm, bt_kq, bt_win, bt_bins;
  int bt_hlo = 0, bt_hr = 0, bt_nbmax = 0;
  double bt_key_fs = -1.0, bt_key_floor = -1.0, bt_key_ceil = -1.0;
  DevBuf bt_codes, bt_inv, bt_hbs, bt_needs, bt_mask, bt_cnt, bt_off, bt_scan, bt_list, bt_cinfo,
      bt_sched, bt_res;
  long long last_nnz = -1;
  long long last_frames = 0;
  bool submitted = false;       // an analysis is enqueued and not yet collected
  long long pending_frames = 0;  // its total frames
  int lds_limit = 64 * 1024;
  // Candidate downloads of the last analysis into h_cand: f0 (cand_ev[0])
  // then scores (cand_ev[1]).  The _async entry points return before they
  // complete; hv_wait_candidates waits for them.
  hipEvent_t cand_ev[2] = {nullptr, nullptr};
  // The candidates are copied into h_cand by the launch sequence itself:
  // f0 at [0, capacity), scores at [capacity, 2 capacity), capacity sized by
  // the last analysis' candidates per frame (an analysis that finds more
  // copies them at collect).
  double cand_per_frame = 20.0;
  long long cand_capacity = 0;  // of the last analysis
  hv_tuning tune{kFftMinHalf, 0, (int)(kSlabBudget >> 20), 1, 0};
  // hv_order_after: an event on another workspace's stream that the next
  // submitted analysis' device work waits for (after its uploads).
  hipEvent_t after_ev = nullptr;
  bool after_pending = false;
  // Executable graphs of the device launch sequence, one per plan shape and
  // buffer set (least recently used first out).
  struct Graph {
    std::vector<long long> key;
    hipGraphExec_t exec = nullptr;
    unsigned long long used = 0;
  };
  std::vector<Graph> graphs;
  unsigned long long graph_clock = 0;
  long long graph_hits = 0, graph_captures = 0;

  explicit hv_workspace(int dev) : device(dev) {
    hipDeviceProp_t prop;
    check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    lds_limit = (int)prop.sharedMemPerBlock;
    check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
    for (auto& e : cand_ev)
      check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreateWithFlags(&after_ev, hipEventDisableTiming), "hipEventCreate");
  }
  ~hv_workspace() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : cand_ev)
      if (e) (void)hipEventDestroy(e);
    if (after_ev) (void)hipEventDestroy(after_ev);
    for (auto& e : timer_events)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace hv {
namespace {

struct Plan {
  Config cfg{};
  std::vector<Utt> utts;
  std::vector<long long> c_off, t_off, f_off, p_off, y_off, x_off;
  long long total_x = 0, total_m = 0, total_y = 0, total_frames = 0;
  long long total_tiles = 0, total_chunks = 0, total_pblocks = 0;
  long long max_chunks = 0;  // largest utterance's IIR chunk count
};

hipStream_t s_of(hv_workspace* ws) { return ws->stream; }

// Two events attributed to timing slot `slot`, for the caller to record
// around a launch on the workspace's stream; a null pair when not timing.
std::pair<hipEvent_t, hipEvent_t> timer_pair(hv_workspace* ws, int slot) {
  if (!ws->timing) return {nullptr, nullptr};
  if (ws->timers_used + 2 > ws->timer_events.size()) {
    for (int k = 0; k < 16; ++k) {
      hipEvent_t e;
      check(hipEventCreate(&e), "hipEventCreate");
      ws->timer_events.push_back(e);
    }
  }
  hipEvent_t a = ws->timer_events[ws->timers_used], b = ws->timer_events[ws->timers_used + 1];
  ws->timers_used += 2;
  ws->timer_slots.push_back(slot);
  return {a, b};
}

// Records `launch` between two events attributed to timing slot `slot`.
template <class F> void timed(hv_workspace* ws, int slot, F&& launch) {
  auto ev = timer_pair(ws, slot);
  if (ev.first) check(hipEventRecord(ev.first, ws->stream), "hipEventRecord");
  launch();
  if (ev.second) check(hipEventRecord(ev.second, ws->stream), "hipEventRecord");
}

void start_timers(hv_workspace* ws, bool on) {
  ws->timing = on;
  ws->timers_used = 0;
  ws->timer_slots.clear();
}

// Sum of the recorded kernel durations per slot (after the stream synced).
void collect_timers(hv_workspace* ws, double* timings) {
  for (size_t k = 0; k < ws->timer_slots.size(); ++k) {
    float ms = 0.f;
    check(hipEventElapsedTime(&ms, ws->timer_events[2 * k], ws->timer_events[2 * k + 1]), "timing");
    timings[ws->timer_slots[k]] += ms;
  }
}

// Band boundaries, half lengths and reversed filters of one (fs_w, floor,
// ceil); cached in the workspace.  GetFilteredSignal harvest.cpp:99-108 and
// NuttallWindow common.cpp:113-121, evaluated with the host libm like WORLD.
void prepare_filters(hv_workspace* ws, double fs_work, double f0_floor,
                     double f0_ceil, const double* given_boundaries, int given_bands) {
  std::vector<double> bounds;
  if (given_boundaries) {
    bounds.assign(given_boundaries, given_boundaries + given_bands);
  } else {
    const double af = f0_floor * 0.9, ac = f0_ceil * 1.1;
    const int bands = 1 + (int)(std::log(ac / af) / kLog2 * 40.0);
    bounds.resize(bands);
    for (int i = 0; i < bands; ++i) bounds[i] = af * std::pow(2.0, (i + 1) / 40.0);
  }
  if (ws->key_fs == fs_work && ws->key_floor == f0_floor && ws->key_ceil == f0_ceil &&
      ws->key_fft_min == ws->tune.fft_min_half && ws->boundaries == bounds)
    return;
  const int B = (int)bounds.size();
  std::vector<int> halves(B), offs(B);
  std::vector<double> coef;
  int max_half = 0;
  for (int b = 0; b < B; ++b) {
    const int h = host_round(fs_work / bounds[b] * 2.0);
    if (h < 1) invalid("band-pass filter shorter than 3 taps; f0_ceil is too high");
    halves[b] = h;
    max_half = std::max(max_half, h);
    const int len = 2 * h + 1, taps8 = (len + 7) & ~7;
    std::vector<double> c(len);
    for (int i = 0; i < len; ++i) {
      const double tmp = i / (len - 1.0);
      c[i] = 0.355768 - 0.487396 * std::cos(2.0 * kPi * tmp) +
             0.144232 * std::cos(4.0 * kPi * tmp) - 0.012604 * std::cos(6.0 * kPi * tmp);
    }
    for (int i = -h; i <= h; ++i) c[i + h] *= std::cos(2 * kPi * bounds[b] * i / fs_work);
    offs[b] = (int)coef.size();
    for (int mm = 0; mm < taps8; ++mm) coef.push_back(mm < len ? c[len - 1 - mm] : 0.0);
  }
  double* dcoef = ws->coef.ensure<double>(coef.size());
  int* doff = ws->coef_off.ensure<int>(B);
  int* dhalf = ws->half.ensure<int>(B);
  double* dbound = ws->bounds.ensure<double>(B);
  check(hipMemcpyAsync(dcoef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice,
                       ws->stream), "upload filters");
  check(hipMemcpyAsync(doff, offs.data(), B * sizeof(int), hipMemcpyHostToDevice, ws->stream),
        "upload filters");
  check(hipMemcpyAsync(dhalf, halves.data(), B * sizeof(int), hipMemcpyHostToDevice, ws->stream),
        "upload filters");
  check(hipMemcpyAsync(dbound, bounds.data(), B * sizeof(double), hipMemcpyHostToDevice,
                       ws->stream), "upload filters");
  // FFT bands: the leading (longest) bands with h >= the tuned minimum
  // (kFftMinHalf by default), when the longest fits a 4096-point block.
  int nf = 0;
  const int fft_min = ws->tune.fft_min_half;
  if (halves[0] <= kFftMaxHalf)
    while (nf < B && halves[nf] >= fft_min && halves[nf] <= halves[0]) ++nf;
  if (nf > 0) {
    if (!ws->tw4096_ready) {
      launch_twiddle(ws->stream, ws->tw4096.ensure<double>(2 * 4096), 4096);
      ws->tw4096_ready = true;
    }
    launch_band_spectra(ws->stream, Config{}, nf, dcoef, doff, dhalf, ws->tw4096.get<double>(),
                        ws->band_spec.ensure<double>((size_t)nf * kSpecDoubles));
  }
  check(hipStreamSynchronize(ws->stream), "upload filters");
  ws->fft_bands = nf;
  ws->fft_D = nf > 0 ? halves[0] : 0;
  ws->key_fs = fs_work;
  ws->key_floor = f0_floor;
  ws->key_ceil = f0_ceil;
  ws->key_fft_min = fft_min;
  ws->boundaries = bounds;
  ws->halves = halves;
  ws->max_half = max_half;
}

// Largest refinement half window: candidates are means of values in
// [f0_floor, f0_ceil]; the 1e-9 slack covers their rounding, the + 1 the
// truncation of harvest.cpp:598.
int refine_hmax(double fs_work, double f0_floor) {
  return (int)(1.5 * fs_work / (f0_floor * (1.0 - 1e-9)) + 1.0) + 1;
}

// The filterbank's phase table for every fft of the plan (grown, cached):
// size P = the next power of two >= max(4096, largest fft), 2^12 fine entries.
void prepare_phase(hv_workspace* ws, Plan& p) {
  constexpr int kFineBits = 12;
  long long need = 1LL << kFineBits;
  for (const Utt& u : p.utts) need = std::max(need, (long long)u.fft);
  long long P = 1LL << kFineBits;
  while (P < need) P <<= 1;
  if (P > ws->phase_n) {
    launch_phase_table(ws->stream,
                       ws->phase.ensure<double>(2 * ((1LL << kFineBits) + (P >> kFineBits))), P,
                       kFineBits);
    ws->phase_n = P;
  }
  p.cfg.phase_n = ws->phase_n;
  p.cfg.phase_bits = kFineBits;
}

// Whether every refinement window depends on its half length alone: at 1 ms
// frames t fs_work is an integer when fs_work is a multiple of 1000 Hz, so
// the Blackman phase offset is -(h + 1) at every frame (k_window_pairs).  The
// table holds (hmax + 1)^2 entries; very low floors keep the per-frame form.
constexpr int kWindowPairsMaxH = 1023;
bool use_window_pairs(double fs_work, int hmax) {
  return fs_work == std::floor(fs_work) && std::fmod(fs_work, 1000.0) == 0.0 &&
         hmax <= kWindowPairsMaxH;
}

void prepare_twiddles(hv_workspace* ws, double fs_work, double f0_floor) {
  const int wh = refine_hmax(fs_work, f0_floor);
  if (wh > ws->wintab_h) {
    double* tab = ws->wintab.ensure<double>(2 * 65 * ((size_t)wh + 1));
    launch_window_table(ws->stream, tab, wh);
    ws->wintab_h = wh;
  }
  if (use_window_pairs(fs_work, wh) && wh > ws->winpair_h) {
    double2* tab = ws->winpair.ensure<double2>(((size_t)wh + 1) * ((size_t)wh + 1));
    launch_window_pairs(ws->stream, tab, wh);
    ws->winpair_h = wh;
  }
  // Largest refinement window: candidates are >= f0_floor (harvest.cpp:250-252).
  const double hmax = 1.5 * fs_work / (f0_floor * (1.0 - 1e-9)) + 2.0;
  int need = 0;
  if (hmax < 1e9) {
    const long long len = 2 * (long long)hmax + 1;
    int lg = 0;
    while ((1LL << (lg + 1)) <= len) ++lg;
    need = lg + 2 <= 20 ? 1 << (lg + 2) : 0;
  }
  if (need > kTwiddleMax) need = 0;
  if (ws->twiddle_n < 0) ws->twiddle_n = 0;
  // A table of size T serves every n dividing T (stride T / n); windows larger
  // than the table fall back to sincospi inside the refinement kernel.
  if (need <= ws->twiddle_n) return;
  double* tw = ws->twiddle.ensure<double>(2 * (size_t)need);
  launch_twiddle(ws->stream, tw, need);
  ws->twiddle_n = need;
}

// Half window of a positive candidate (harvest.cpp:596), as the kernels
// evaluate it.
int half_window_of(double fs_work, double f0) { return (int)(1.5 * fs_work / f0 + 1.0); }

// Smallest double in [lo, hi] (both positive) where `pred` holds, for a
// predicate false below some point and true from it on; hi when it holds
// nowhere below hi.  Positive doubles order like their bit patterns.
template <class P> double first_double(double lo, double hi, P pred) {
  uint64_t a, b;
  std::memcpy(&a, &lo, sizeof a);
  std::memcpy(&b, &hi, sizeof b);
  while (a < b) {
    const uint64_t mid = a + (b - a) / 2;
    double m;
    std::memcpy(&m, &mid, sizeof m);
    if (pred(m)) b = mid;
    else a = mid + 1;
  }
  double r;
  std::memcpy(&r, &a, sizeof r);
  return r;
}

// Bin tables of the refinement for one (fs_work, floor, ceil), cached in
// the workspace (ws->bt_hr = 0: the per-frame refinement serves the key).
// Candidates lie in [f0_floor, f0_ceil] up to rounding (means of values the
// band gate keeps, harvest.cpp:250-252), so their half windows span
// [half(ceil), refine_hmax].  The candidates of one half window h form the
// doubles interval [flo, fhi] where half_window_of is h (it is monotone in
// f0), and FixF0's bin of harmonic j, round(f0 n / fs j) (harvest.cpp:515),
// is monotone in f0 too: the bins of h and j are exactly
// [round(flo n / fs j), round(fhi n / fs j)], for j up to
// min(fs / 2 / flo, 6) (harvest.cpp:577-578).  The kernels check every
// slot against these ranges all the same.
void prepare_bintable(hv_workspace* ws, double fs_work, double f0_floor, double f0_ceil) {
  if (ws->bt_key_fs == fs_work && ws->bt_key_floor == f0_floor && ws->bt_key_ceil == f0_ceil)
    return;
  ws->bt_hr = 0;
  ws->bt_key_fs = fs_work;
  ws->bt_key_floor = f0_floor;
  ws->bt_key_ceil = f0_ceil;
  const int hhi = refine_hmax(fs_work, f0_floor);
  if (!use_window_pairs(fs_work, hhi)) return;
  const int hlo = half_window_of(fs_work, f0_ceil * (1.0 + 1e-9));
  const int hr = hhi - hlo + 1;
  if (hlo < 1 || hr >= kBtMaxHr) return;  // hb < 0xFF: a byte per candidate, 0xFF for none
  const double lo = f0_floor * 0.5, hi = f0_ceil * 2.0;  // beyond every candidate
  std::vector<int4> info(hr), harm(6 * (size_t)hr);
  std::vector<std::vector<int>> bins(hr);
  long long rows = 0, records = 0;
  int nbmax = 1;
  for (int hb = 0; hb < hr; ++hb) {
    const int h = hlo + hb, n = window_fft(h);
    const double flo = first_double(lo, hi, [&](double f) { return half_window_of(fs_work, f) <= h; });
    const double fnext = first_double(lo, hi, [&](double f) { return half_window_of(fs_work, f) < h; });
    const double fhi = half_window_of(fs_work, hi) < h ? std::nextafter(fnext, 0.0) : hi;
    const bool empty = half_window_of(fs_work, flo) != h || fhi < flo;
    const int H = empty ? 0 : std::min((int)(fs_work / 2.0 / flo), 6);
    for (int j = 0; j < 6; ++j) {
      int4 e{0, 0, (int)bins[hb].size(), 0};
      if (j < H) {
        const int klo = mround(flo * n / fs_work * (j + 1));
        const int khi = mround(fhi * n / fs_work * (j + 1));
        e.x = klo;
        e.y = khi - klo + 1;
        for (int k = klo; k <= khi; ++k) bins[hb].push_back(k);
      }
      harm[6 * hb + j] = e;
    }
    const int nb = (int)bins[hb].size();
    if (nb > 127) return;  // candidate codes hold 7-bit indices of the table bins
    info[hb] = int4{n, nb, (int)rows, (int)records};
    rows += 2 * h + 1;
    records += nb;
    nbmax = std::max(nbmax, nb);
  }
  std::vector<int> kq((size_t)hr * nbmax, 0);
  for (int hb = 0; hb < hr; ++hb)
    for (size_t q = 0; q < bins[hb].size(); ++q) kq[(size_t)hb * nbmax + q] = bins[hb][q];
  int4* dinfo = ws->bt_info.ensure<int4>(hr);
  int4* dharm = ws->bt_harm.ensure<int4>(6 * (size_t)hr);
  int* dkq = ws->bt_kq.ensure<int>(kq.size());
  double2* dwin = ws->bt_win.ensure<double2>((size_t)rows);
  double* dbins = ws->bt_bins.ensure<double>((size_t)std::max(1LL, records) * kBtRecord);
  check(hipMemcpyAsync(dinfo, info.data(), info.size() * sizeof(int4), hipMemcpyHostToDevice,
                       ws->stream), "upload bin tables");
  check(hipMemcpyAsync(dharm, harm.data(), harm.size() * sizeof(int4), hipMemcpyHostToDevice,
                       ws->stream), "upload bin tables");
  check(hipMemcpyAsync(dkq, kq.data(), kq.size() * sizeof(int), hipMemcpyHostToDevice,
                       ws->stream), "upload bin tables");
  Config c{};
  c.fs_work = fs_work;
  c.bt_hlo = hlo;
  c.bt_hr = hr;
  c.bt_nbmax = nbmax;
  launch_bt_tables(ws->stream, c, BinTable{dinfo, dharm, dwin, dbins}, dkq, dwin, dbins);
  check(hipStreamSynchronize(ws->stream), "bin tables");
  ws->bt_hlo = hlo;
  ws->bt_hr = hr;
  ws->bt_nbmax = nbmax;
}

// The refinement path of an analysis: the bin tables when the workspace's
// tuning allows them and they exist for the analysis' key.
void apply_bintable(hv_workspace* ws, Config& c) {
  const bool on = ws->tune.refine_table && ws->bt_hr > 0 && ws->bt_key_fs == c.fs_work &&
                  ws->bt_key_floor == c.f0_floor && ws->bt_key_ceil == c.f0_ceil;
  c.bt_hlo = on ? ws->bt_hlo : 0;
  c.bt_hr = on ? ws->bt_hr : 0;
  c.bt_nbmax = on ? ws->bt_nbmax : 0;
  c.bt_groups = ws->tune.refine_groups > 0 ? ws->tune.refine_groups : kBtDftGroups;
}

void fill_config(Config& c, int count, int ratio, int fs, double fs_work, double f0_floor,
                 double f0_ceil, int bands, int capacity, int twiddle_n) {
  c.count = count;
  c.ratio = ratio;
  c.lag = ratio > 1 ? (int)(std::ceil(140.0 / ratio) * ratio) : 0;
  c.fs = fs;
  c.fs_work = fs_work;
  c.f0_floor = f0_floor;
  c.f0_ceil = f0_ceil;
  c.bands = bands;
  c.capacity = capacity;
  c.slots = 7 * capacity;
  c.twiddle_n = twiddle_n;
  c.lds_limit = 64 * 1024;
  c.hmax = refine_hmax(fs_work, f0_floor);
  c.win_pairs = use_window_pairs(fs_work, c.hmax) ? 1 : 0;
  c.bt_hlo = c.bt_hr = c.bt_nbmax = 0;  // apply_bintable turns the bin tables on
  c.bt_groups = kBtDftGroups;
  const double* co = kDecimation[ratio >= 2 && ratio <= 12 ? ratio : 0];
  c.a0 = co[0];
  c.a1 = co[1];
  c.a2 = co[2];
  c.b0 = co[3];
  c.b1 = co[4];
  chunk_power(co, c.chunk_pow);
}

// Per-utterance layout.  `ns`, `ffts` and `frames` override WORLD's formulas
// for the y-input entry point (hv_analyze).
void layout(Plan& p, const std::vector<long long>& lengths, const std::vector<int>& ns,
            const std::vector<int>& ffts, const std::vector<int>& frames) {
  const int U = (int)lengths.size();
  const Config& c = p.cfg;
  p.utts.resize(U);
  p.c_off.assign(U + 1, 0);
  p.t_off.assign(U + 1, 0);
  p.f_off.assign(U + 1, 0);
  p.p_off.assign(U + 1, 0);
  p.y_off.assign(U + 1, 0);
  p.x_off.assign(U + 1, 0);
  long long x = 0, m = 0;
  for (int u = 0; u < U; ++u) {
    Utt& t = p.utts[u];
    std::memset(&t, 0, sizeof(t));
    t.x_off = x;
    t.length = lengths[u];
    t.m_off = m;
    if (c.ratio > 1) {
      const long long padded = t.length + 2LL * c.lag;
      t.m = padded + 18;
      const long long nout = (padded - 1) / c.ratio + 1;
      t.nbeg = (int)(c.ratio - c.ratio * nout + padded);
    }
    t.n = ns[u];
    t.fft = ffts[u];
    t.frames = frames[u];
    t.tiles = (int)((t.n + kTileStride - 1) / kTileStride);
    t.chunks = (int)((t.m + kChunk - 1) / kChunk);
    t.pblocks = (int)((t.n + kMeanBlock - 1) / kMeanBlock);
    t.y_off = p.y_off[u];
    t.f_off = p.f_off[u];
    t.t_off = p.t_off[u];
    t.c_off = p.c_off[u];
    t.p_off = p.p_off[u];
    p.y_off[u + 1] = p.y_off[u] + t.n;
    p.f_off[u + 1] = p.f_off[u] + t.frames;
    p.t_off[u + 1] = p.t_off[u] + t.tiles;
    p.c_off[u + 1] = p.c_off[u] + t.chunks;
    p.max_chunks = std::max(p.max_chunks, (long long)t.chunks);
    p.p_off[u + 1] = p.p_off[u] + t.pblocks;
    x += t.length;
    m += t.m;
    p.x_off[u + 1] = x;
  }
  p.total_x = x;
  p.total_m = m;
  p.total_y = p.y_off[U];
  p.total_frames = p.f_off[U];
  p.total_tiles = p.t_off[U];
  p.total_chunks = p.c_off[U];
  p.total_pblocks = p.p_off[U];
  if (p.total_frames * (long long)c.slots >= (1LL << 31) - 1)
    invalid("input or analysis settings exceed supported allocation sizes");
}

// Bytes of a plan's descriptor arrays: the utterances, then six offset
// arrays of U + 1 entries.
constexpr int kPlanArrays = 6;
size_t plan_bytes(int U) { return U * sizeof(Utt) + kPlanArrays * (U + 1) * sizeof(long long); }

// Upload the plan's descriptor arrays through one pinned staging buffer.
void upload_plan(hv_workspace* ws, const Plan& p) {
  const int U = p.cfg.count;
  constexpr int kArrays = kPlanArrays;
  const size_t bytes = plan_bytes(U);
  char* h = ws->h_meta.ensure<char>(bytes);
  std::memcpy(h, p.utts.data(), U * sizeof(Utt));
  size_t at = U * sizeof(Utt);
  const std::vector<long long>* arrays[kArrays] = {&p.c_off, &p.t_off, &p.f_off,
                                                   &p.p_off, &p.y_off, &p.x_off};
  const long long** views[kArrays] = {&ws->d_c_off, &ws->d_t_off, &ws->d_f_off,
                                      &ws->d_p_off, &ws->d_y_off, &ws->d_x_off};
  char* d = ws->plan.ensure<char>(bytes);
  ws->d_utt = reinterpret_cast<const Utt*>(d);
  for (int k = 0; k < kArrays; ++k) {
    std::memcpy(h + at, arrays[k]->data(), (U + 1) * sizeof(long long));
    *views[k] = reinterpret_cast<const long long*>(d + at);
    at += (U + 1) * sizeof(long long);
  }
  check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ws->stream), "upload plan");
}

// Device buffers of one analysis, all sized before anything is enqueued (so
// that a captured launch sequence never allocates).
struct Bufs {
  const double* dx = nullptr;  // input: the caller's device batch or ws->x
  double *y, *pass, *states, *carries, *carry_tmp, *partial, *mean;
  Quirk* quirk;
  double *partial3, *sums3;
  int group;  // bands per slab group
  double* slab;
  int *counts, *offs, *totals;
  double *raw, *tspec;
  int* count;
  double *official, *rf0, *rsc, *pf0, *psc, *dbest, *cinit;
  unsigned long long* work;
  int *row_nnz, *doffsets, *scan_tmp;
  long long nb;
  int32_t* hoff;
  double* hbest;
  unsigned long long* hwork;
  long long capacity;  // candidates the sequence copies into h_cand (0: none)
  double* hcand;       // h_cand when capacity > 0
  BtBufs bt;           // bin-table refinement (when cfg.bt_hr > 0)
};

// HV_STAT_SLOTS work counters, then a flag the refinement sets if a
// candidate's window exceeds the bound its LDS was sized for; each spread
// over kWorkSpread addresses.
constexpr int kWorkWords = (HV_STAT_SLOTS + 1) * kWorkSpread;
static_assert(kStatFault == HV_STAT_SLOTS && kStatTablePairs < HV_STAT_SLOTS,
              "device counter layout must match HV_STAT_SLOTS");

Bufs size_buffers(hv_workspace* ws, const Plan& p, const double* dx) {
  const Config& c = p.cfg;
  const int U = c.count, B = c.bands;
  Bufs b{};
  b.dx = dx;
  b.y = ws->y.ensure<double>(p.total_y);
  b.pass = ws->pass.ensure<double>(p.total_m);
  b.states = ws->states.ensure<double>(3 * p.total_chunks);
  b.carries = ws->carries.ensure<double>(3 * p.total_chunks);
  b.carry_tmp = ws->carry_tmp.ensure<double>(carry_scratch_doubles(U, p.max_chunks));
  b.partial = ws->partial.ensure<double>(p.total_pblocks);
  b.mean = ws->mean.ensure<double>(U);
  b.quirk = ws->quirk.ensure<Quirk>((size_t)U * B);
  b.partial3 = ws->partial3.ensure<double>(3 * p.total_pblocks);
  b.sums3 = ws->sums3.ensure<double>(3 * U);
  // Band groups bounded by the slab budget.
  const size_t budget = (size_t)ws->tune.slab_budget_mib << 20;
  const size_t per_band = (size_t)4 * p.total_tiles * kSlabCap * sizeof(double);
  b.group = (int)std::max<size_t>(1, std::min<size_t>(B, budget / std::max<size_t>(per_band, 1)));
  b.slab = ws->slab.ensure<double>((size_t)b.group * 4 * p.total_tiles * kSlabCap);
  b.counts = ws->counts.ensure<int>((size_t)b.group * 4 * p.total_tiles);
  b.offs = ws->offs.ensure<int>((size_t)b.group * 4 * p.total_tiles);
  b.totals = ws->totals.ensure<int>((size_t)b.group * 4 * U);
  b.raw = ws->raw.ensure<double>((size_t)B * p.total_frames);
  b.tspec = ws->fft_bands > 0 ? ws->tile_spec.ensure<double>((size_t)p.total_tiles * kSpecDoubles)
                              : nullptr;
  b.count = ws->count.ensure<int>(U);
  b.official = ws->official.ensure<double>((size_t)p.total_frames * c.capacity);
  const size_t slots = (size_t)p.total_frames * c.slots;
  b.rf0 = ws->rf0.ensure<double>(slots);
  b.rsc = ws->rsc.ensure<double>(slots);
  b.work = ws->work.ensure<unsigned long long>(kWorkWords);
  b.pf0 = ws->pf0.ensure<double>(slots);
  b.psc = ws->psc.ensure<double>(slots);
  b.dbest = ws->best.ensure<double>(p.total_frames);
  b.row_nnz = ws->row_nnz.ensure<int>(p.total_frames);
  b.doffsets = ws->offsets.ensure<int>(p.total_frames + 1);
  b.nb = scan_block_count(p.total_frames);
  b.scan_tmp = ws->scan_tmp.ensure<int>(b.nb + 1);
  b.cinit = ws->cinit.ensure<double>(slots);
  b.hoff = ws->h_offsets.ensure<int32_t>(p.total_frames + 1);
  b.hbest = ws->h_best.ensure<double>(p.total_frames);
  b.hwork = ws->h_work.ensure<unsigned long long>(kWorkWords);
  b.capacity = 0;
  b.hcand = nullptr;
  b.bt = BtBufs{};
  if (c.bt_hr > 0) {
    const long long nblk = bt_blocks(p.total_frames), cells = (long long)c.bt_hr * nblk;
    b.bt.codes = ws->bt_codes.ensure<unsigned long long>((size_t)p.total_frames * c.capacity);
    b.bt.inv = ws->bt_inv.ensure<double>((size_t)p.total_frames * c.capacity);
    b.bt.hbs = ws->bt_hbs.ensure<unsigned char>((size_t)p.total_frames * bt_hb_stride(c.capacity));
    b.bt.needs = ws->bt_needs.ensure<unsigned long long>((size_t)p.total_frames * kBtWords);
    b.bt.mask = ws->bt_mask.ensure<unsigned long long>(cells);
    b.bt.cnt = ws->bt_cnt.ensure<int>(cells);
    b.bt.off = ws->bt_off.ensure<int>(cells + 1);
    b.bt.scan_len = scan_block_count(cells) + 1;
    b.bt.scan_tmp = ws->bt_scan.ensure<int>(b.bt.scan_len);
    b.bt.list = ws->bt_list.ensure<int2>((size_t)p.total_frames * kBtPairs);
    b.bt.cinfo = ws->bt_cinfo.ensure<int>(3 * kBtMaxHr + 1);
    b.bt.sched = ws->bt_sched.ensure<int>(1);
    b.bt.res = ws->bt_res.ensure<double2>((size_t)p.total_frames * kBtPairs * c.bt_nbmax);
  }
  return b;
}

// Sizes the in-sequence candidate copy of a two-step analysis.
void size_candidate_copy(hv_workspace* ws, const Plan& p, Bufs& b) {
  const long long slots = p.total_frames * (long long)p.cfg.slots;
  b.capacity = std::min<long long>(slots, std::max<long long>(
      1024, (long long)std::ceil(ws->cand_per_frame * (double)p.total_frames)));
  b.hcand = ws->h_cand.ensure<double>(2 * (size_t)b.capacity);
}

// Decimation (or the copy at ratio 1) and mean removal: y on the device.
void enqueue_preprocess(hv_workspace* ws, const Plan& p, const Bufs& b, bool events) {
  hipStream_t s = ws->stream;
  timed(ws, 3, [&] {
    if (p.cfg.ratio > 1)
      launch_decimate(s, p.cfg, ws->d_utt, ws->d_c_off, p.total_chunks, p.max_chunks, b.dx,
                      b.pass, b.states, b.carries, b.carry_tmp, b.y);
    else
      launch_copy_input(s, p.cfg, ws->d_utt, ws->d_y_off, p.total_y, b.dx, b.y);
  });
  launch_remove_mean(s, p.cfg, ws->d_utt, ws->d_p_off, p.total_pblocks, ws->d_y_off, p.total_y,
                     b.y, b.partial, b.mean);
  if (events) check(hipEventRecord(ws->ev[1], s), "hipEventRecord");
}

// Everything after the decimated, mean-removed y, through the download of
// offsets, best and the work counters; the host does not wait.
void enqueue_analysis(hv_workspace* ws, const Plan& p, const Bufs& b, bool events) {
  const Config& c = p.cfg;
  const int U = c.count, B = c.bands;
  hipStream_t s = ws->stream;
  const Utt* dutt = ws->d_utt;
  const double* coef = ws->coef.get<double>();
  const int* coef_off = ws->coef_off.get<int>();
  const int* half = ws->half.get<int>();
  launch_quirk(s, c, dutt, ws->d_p_off, p.total_pblocks, b.y, coef, coef_off, half, b.partial3,
               b.sums3, b.quirk);
  // The long bands [0, nfft) by FFT (one spectrum per tile, shared by the
  // bands), the others by the direct FIR.
  const int nfft = ws->fft_bands, lead = ws->fft_D;
  if (nfft > 0) {
    timed(ws, 4, [&] {
      launch_tile_spectra(s, c, dutt, ws->d_t_off, p.total_tiles, lead, b.y,
                          ws->tw4096.get<double>(), b.tspec);
    });
  }
  for (int b0 = 0; b0 < B; b0 += b.group) {
    const int g = std::min(b.group, B - b0);
    const int nf = std::max(0, std::min(g, nfft - b0));
    int mh = 0;
    for (int k = b0 + nf; k < b0 + g; ++k) mh = std::max(mh, ws->halves[k]);
    timed(ws, 4, [&] {
      launch_filter_fft(s, c, dutt, ws->d_t_off, p.total_tiles, b0, nf, 0, lead, b.tspec,
                        ws->band_spec.get<double>(), half, b.quirk, ws->tw4096.get<double>(),
                        ws->phase.get<double>(), b.slab, b.counts);
      launch_filter_events(s, c, dutt, ws->d_t_off, p.total_tiles, b0 + nf, g - nf, nf, mh, b.y,
                           coef, coef_off, half, b.quirk, ws->phase.get<double>(), b.slab,
                           b.counts);
    });
    timed(ws, 5, [&] {
      launch_tile_scan(s, c, dutt, p.total_tiles, g, b.counts, b.offs, b.totals);
      launch_interp(s, c, dutt, ws->d_t_off, p.total_tiles, p.total_frames, b0, g,
                    ws->bounds.get<double>(), b.slab, b.counts, b.offs, b.totals, b.raw);
    });
  }
  launch_fill_int(s, b.count, U, 0);
  launch_official(s, c, ws->d_f_off, p.total_frames, b.raw, b.official, b.count);
  if (events) check(hipEventRecord(ws->ev[2], s), "hipEventRecord");
  launch_fill_int(s, reinterpret_cast<int*>(b.work), 2 * kWorkWords, 0);
  timed(ws, 6, [&] {
    if (c.bt_hr > 0) {
      auto dft = timer_pair(ws, 8);
      launch_refine_table(s, c, dutt, ws->d_f_off, p.total_frames, b.y, b.official, b.count,
                          BinTable{ws->bt_info.get<int4>(), ws->bt_harm.get<int4>(),
                                   ws->bt_win.get<double2>(), ws->bt_bins.get<double>()},
                          b.bt, b.rf0, b.rsc, b.work, dft.first, dft.second);
    } else
      launch_refine(s, c, dutt, ws->d_f_off, p.total_frames, b.y, b.official, b.count,
                    ws->twiddle.get<double>(), ws->wintab.get<double>(),
                    c.win_pairs ? ws->winpair.get<double2>() : nullptr, b.rf0, b.rsc, b.work);
  });
  timed(ws, 7, [&] {
    launch_prune_best(s, c, dutt, ws->d_f_off, p.total_frames, b.count, b.rf0, b.rsc, b.pf0,
                      b.psc, b.dbest, b.row_nnz);
    launch_exclusive_scan(s, b.row_nnz, b.doffsets, p.total_frames, b.scan_tmp, b.nb + 1);
    // Compacted candidates reuse the refinement buffers (no longer needed).
    launch_compact(s, c, dutt, ws->d_f_off, p.total_frames, b.count, b.official, b.pf0, b.psc,
                   b.doffsets, b.rf0, b.rsc, b.cinit);
  });
  // Two-step (pipelined) analyses copy the candidates into pinned host
  // memory in the same sequence (the first `capacity` of each array; the
  // device's total is known only at collect): the host finds them in place
  // when the sequence ends, with no download queued behind the other
  // workspace's kernels.  Direct calls download them at collect instead,
  // overlapping the first contour steps.
  if (b.capacity > 0) {
    check(hipMemcpyAsync(b.hcand, b.rf0, b.capacity * sizeof(double), hipMemcpyDeviceToHost, s),
          "download candidates");
    check(hipMemcpyAsync(b.hcand + b.capacity, b.rsc, b.capacity * sizeof(double),
                         hipMemcpyDeviceToHost, s), "download candidates");
  }
  check(hipMemcpyAsync(b.hoff, b.doffsets, (p.total_frames + 1) * sizeof(int32_t),
                       hipMemcpyDeviceToHost, s), "download offsets");
  check(hipMemcpyAsync(b.hbest, b.dbest, p.total_frames * sizeof(double), hipMemcpyDeviceToHost,
                       s), "download best");
  check(hipMemcpyAsync(b.hwork, b.work, kWorkWords * sizeof(unsigned long long),
                       hipMemcpyDeviceToHost, s), "download counters");
}

// What a captured launch sequence depends on: every size and pointer its
// kernels and copies take, and the settings folded into the launch config.
std::vector<long long> graph_key(hv_workspace* ws, const Plan& p, const Bufs& b) {
  const Config& c = p.cfg;
  std::vector<long long> k = {c.count, c.ratio, c.fs, c.bands, c.capacity, c.twiddle_n,
                              c.hmax, (long long)c.phase_n, ws->fft_bands, ws->fft_D, b.group,
                              c.bt_hlo, c.bt_hr, c.bt_nbmax, c.bt_groups,
                              p.total_x, p.total_m, p.total_y, p.total_frames, p.total_tiles,
                              p.total_chunks, p.total_pblocks, p.max_chunks};
  long long bits;
  for (double v : {c.fs_work, c.f0_floor, c.f0_ceil}) {
    std::memcpy(&bits, &v, sizeof(bits));
    k.push_back(bits);
  }
  const void* ptrs[] = {b.dx, ws->d_utt, ws->coef.p, ws->coef_off.p, ws->half.p, ws->bounds.p,
                        ws->twiddle.p, ws->wintab.p, ws->winpair.p, ws->tw4096.p, ws->band_spec.p, ws->phase.p,
                        b.y, b.pass, b.states, b.carries, b.carry_tmp, b.partial, b.mean, b.quirk,
                        b.partial3, b.sums3, b.slab, b.counts, b.offs, b.totals, b.raw, b.tspec,
                        b.count, b.official, b.rf0, b.rsc, b.pf0, b.psc, b.dbest, b.cinit, b.work,
                        b.row_nnz, b.doffsets, b.scan_tmp, b.hoff, b.hbest, b.hwork,
                        b.hcand, ws->bt_info.p, ws->bt_harm.p, ws->bt_win.p, ws->bt_bins.p,
                        b.bt.codes, b.bt.inv, b.bt.hbs, b.bt.needs,
                        b.bt.mask, b.bt.cnt, b.bt.off, b.bt.scan_tmp, b.bt.list, b.bt.cinfo,
                        b.bt.sched, b.bt.res};
  for (const void* q : ptrs) k.push_back((long long)(uintptr_t)q);
  k.push_back(b.capacity);
  return k;
}

constexpr size_t kMaxGraphs = 8;

// The device launch sequence as one graph launch: replayed when a graph of
// the same key exists, else captured from the stream (nothing in it
// allocates or clears memory with a memset: counters are cleared by fill
// kernels), instantiated and cached.
void launch_as_graph(hv_workspace* ws, const Plan& p, const Bufs& b) {
  hipStream_t s = ws->stream;
  std::vector<long long> key = graph_key(ws, p, b);
  ++ws->graph_clock;
  for (auto& g : ws->graphs)
    if (g.key == key) {
      g.used = ws->graph_clock;
      ++ws->graph_hits;
      check(hipGraphLaunch(g.exec, s), "hipGraphLaunch");
      return;
    }
  hipGraph_t graph = nullptr;
  check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  try {
    enqueue_preprocess(ws, p, b, false);
    enqueue_analysis(ws, p, b, false);
  } catch (...) {
    (void)hipStreamEndCapture(s, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  check(hipStreamEndCapture(s, &graph), "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  check(e, "hipGraphInstantiate");
  if (ws->graphs.size() >= kMaxGraphs) {
    auto lru = std::min_element(ws->graphs.begin(), ws->graphs.end(),
                                [](const auto& x, const auto& y) { return x.used < y.used; });
    (void)hipGraphExecDestroy(lru->exec);
    ws->graphs.erase(lru);
  }
  ws->graphs.push_back({std::move(key), exec, ws->graph_clock});
  ++ws->graph_captures;
  check(hipGraphLaunch(exec, s), "hipGraphLaunch");
}

// Enqueues the device work of a staged batch (see stage_batch) and marks the
// workspace submitted: as a graph when graphs are on and no timings were
// asked for, else launch by launch with span events.
// With dx == nullptr, y is already in place (hv_analyze): no preprocessing,
// and the sequence runs launch by launch.
void analyze_submit(hv_workspace* ws, const Plan& p, const double* dx, bool copy_candidates) {
  Bufs b = size_buffers(ws, p, dx);
  if (copy_candidates) size_candidate_copy(ws, p, b);
  if (dx && ws->tune.graphs && !ws->timing) {
    launch_as_graph(ws, p, b);
  } else {
    if (dx) enqueue_preprocess(ws, p, b, true);
    enqueue_analysis(ws, p, b, true);
  }
  check(hipEventRecord(ws->ev[3], s_of(ws)), "hipEventRecord");
  ws->cand_capacity = b.capacity;
  ws->pending_frames = p.total_frames;
  ws->submitted = true;
}

// Waits for the submitted analysis, checks its fault flag, starts the
// candidate download (waited for here unless `defer`) and returns offsets,
// best and nnz.
void analyze_collect(hv_workspace* ws, int32_t* offsets, double* best, int64_t* nnz,
                     double* timings, bool defer) {
  if (!ws->submitted) invalid("no submitted analysis to collect");
  ws->submitted = false;
  const long long frames = ws->pending_frames;
  hipStream_t s = ws->stream;
  const int32_t* hoff = ws->h_offsets.get<int32_t>();
  const double* hbest = ws->h_best.get<double>();
  unsigned long long* hwork = ws->h_work.get<unsigned long long>();
  const double* rf0 = ws->rf0.get<double>();
  const double* rsc = ws->rsc.get<double>();
  check(hipStreamSynchronize(s), "analysis");
  for (int k = 0; k <= HV_STAT_SLOTS; ++k) {  // fold to one entry per counter
    unsigned long long total = 0;
    for (int q = 0; q < kWorkSpread; ++q) total += hwork[k * kWorkSpread + q];
    hwork[k] = total;
  }
  if (hwork[HV_STAT_SLOTS])
    throw std::runtime_error("refinement window beyond its precomputed bound");
  // The candidates end in h_cand (f0 at [0, capacity), scores at
  // [capacity, ...)), where the host contour reads them in place
  // (hv_result_view) or callers copy them out.
  const long long n = hoff[frames];
  if (ws->cand_capacity > 0 && n <= ws->cand_capacity) {
    // Copied by the sequence: both parts are in place already.
    check(hipEventRecord(ws->cand_ev[0], s), "hipEventRecord");
    check(hipEventRecord(ws->cand_ev[1], s), "hipEventRecord");
  } else {
    // Downloaded now, f0 then scores (an in-sequence copy that was too small
    // is redone in full).  The stream is idle, so nothing else targets h_cand.
    const bool overflow = ws->cand_capacity > 0;
    const long long cap = std::max(n, 1LL);
    double* hc = ws->h_cand.ensure<double>(2 * (size_t)cap);
    if (n > 0)
      check(hipMemcpyAsync(hc, rf0, n * sizeof(double), hipMemcpyDeviceToHost, s),
            "download candidates");
    check(hipEventRecord(ws->cand_ev[0], s), "hipEventRecord");
    if (n > 0)
      check(hipMemcpyAsync(hc + cap, rsc, n * sizeof(double), hipMemcpyDeviceToHost, s),
            "download candidates");
    check(hipEventRecord(ws->cand_ev[1], s), "hipEventRecord");
    ws->cand_capacity = cap;
    if (!defer || overflow) check(hipStreamSynchronize(s), "download candidates");
  }
  // The next in-sequence copy takes 10 % more than this analysis' density
  // (at least 20 candidates per frame; the usual density is 15-19).
  ws->cand_per_frame = std::max(20.0, 1.1 * (double)n / (double)std::max(frames, 1LL));
  if (offsets) std::memcpy(offsets, hoff, (frames + 1) * sizeof(int32_t));
  if (best) std::memcpy(best, hbest, frames * sizeof(double));
  ws->last_nnz = n;
  ws->last_frames = frames;
  *nnz = ws->last_nnz;
  if (timings) {
    float ms[3];
    check(hipEventElapsedTime(&ms[0], ws->ev[0], ws->ev[1]), "timing");
    check(hipEventElapsedTime(&ms[1], ws->ev[1], ws->ev[2]), "timing");
    check(hipEventElapsedTime(&ms[2], ws->ev[2], ws->ev[3]), "timing");
    for (int k = 0; k < HV_TIMING_SLOTS; ++k) timings[k] = k < 3 ? ms[k] : 0.0;
    collect_timers(ws, timings);
  }
  ws->timing = false;
}

void analyze_device(hv_workspace* ws, const Plan& p, const double* dx, int32_t* offsets,
                    double* best, int64_t* nnz, double* timings, bool defer = false) {
  analyze_submit(ws, p, dx, false);
  analyze_collect(ws, offsets, best, nnz, timings, defer);
}

template <class F> hv_status guarded(F&& body) {
  try {
    body();
    g_error.clear();
    return HV_OK;
  } catch (const Failure& f) {
    g_error = f.what();
    return f.status;
  } catch (const std::bad_alloc&) {
    g_error = "host memory allocation failed";
    return HV_ENOMEM;
  } catch (const std::exception& e) {
    g_error = e.what();
    return HV_EINTERNAL;
  } catch (...) {
    g_error = "unknown error";
    return HV_EINTERNAL;
  }
}

// WORLD's per-utterance sizes (harvest.cpp:1149-1181, 1219-1221).
void world_sizes(long long length, int fs, double f0_floor, int ratio, double fs_work, int* n,
                 int* fft, int* frames) {
  *n = (int)std::ceil((double)length / ratio);
  const double b0 = f0_floor * 0.9 * std::pow(2.0, 1 / 40.0);
  *fft = suitable_fft_size(*n + 5 + 2 * (int)(2.0 * fs_work / b0));
  *frames = (int)(1000.0 * (double)length / fs / 1.0) + 1;
}

// Host input is staged and uploaded in slices of this many samples (8 MiB).
constexpr long long kUploadSlice = 1LL << 20;

void check_common(int fs, double f0_floor, double f0_ceil) {
  if (fs < 8000 || fs > 192000) invalid("fs must be between 8000 and 192000 Hz");
  if (!std::isfinite(f0_floor) || !std::isfinite(f0_ceil) || f0_floor <= 0 || f0_ceil <= f0_floor)
    invalid("require finite 0 < f0_floor < f0_ceil");
}

// Shared body of the waveform entry points; `device_x` is true when xs[0]
// already points to the concatenated batch in device memory.
// Plan, upload and preprocessing shared by the analysis and the frozen-decision
// entry point: on return the decimated, mean-removed y is on the device.
// Validation, the per-rate tables and the layout of a batch.
Plan plan_batch(hv_workspace* ws, int count, const int64_t* lengths, int fs, double f0_floor,
                double f0_ceil) {
  if (count < 1) invalid("count must be positive");
  check_common(fs, f0_floor, f0_ceil);
  const int ratio = std::max(std::min(host_round(fs / 8000.0), 12), 1);
  const double fs_work = (double)fs / ratio;
  if (f0_ceil >= fs_work / 2) invalid("f0_ceil must be below the internal Nyquist frequency");
  std::vector<long long> lens(count);
  std::vector<int> ns(count), ffts(count), frames(count);
  for (int u = 0; u < count; ++u) {
    if (lengths[u] < 2LL * ratio || lengths[u] > 0x7fffffffLL)
      invalid("x is too short: require at least two samples after internal decimation");
    lens[u] = lengths[u];
    world_sizes(lens[u], fs, f0_floor, ratio, fs_work, &ns[u], &ffts[u], &frames[u]);
  }
  prepare_filters(ws, fs_work, f0_floor, f0_ceil, nullptr, 0);
  prepare_twiddles(ws, fs_work, f0_floor);
  prepare_bintable(ws, fs_work, f0_floor, f0_ceil);
  Plan p;
  const int B = (int)ws->boundaries.size();
  fill_config(p.cfg, count, ratio, fs, fs_work, f0_floor, f0_ceil, B, host_round(B / 10.0),
              ws->twiddle_n);
  apply_bintable(ws, p.cfg);
  p.cfg.lds_limit = ws->lds_limit;
  layout(p, lens, ns, ffts, frames);
  prepare_phase(ws, p);
  return p;
}

Plan stage_batch(hv_workspace* ws, int count, const double* const* xs, const int64_t* lengths,
                 int fs, double f0_floor, double f0_ceil, double* timings, bool device_x,
                 const double** staged) {
  Plan p = plan_batch(ws, count, lengths, fs, f0_floor, f0_ceil);
  hipStream_t s = ws->stream;
  start_timers(ws, timings != nullptr);
  check(hipEventRecord(ws->ev[0], s), "hipEventRecord");
  upload_plan(ws, p);
  const double* dx;
  if (device_x) {
    dx = xs[0];
  } else {
    // The batch goes through pinned staging in slices: the copy of slice k+1
    // (which also checks that every sample is finite; parts of it on several
    // host threads) overlaps the upload of slice k.  Slices split utterances
    // freely; each utterance is contiguous at its x_off in both buffers.
    double* hx = ws->h_x.ensure<double>(p.total_x);
    double* d = ws->x.ensure<double>(p.total_x);
    bool finite = true;
    int u = 0;
    std::vector<CopyPiece> pieces;
    for (long long at = 0; at < p.total_x; at += kUploadSlice) {
      const long long end = std::min(at + kUploadSlice, p.total_x);
      pieces.clear();
      for (; u < count && p.utts[u].x_off < end; ++u) {
        const long long a = std::max(at, p.utts[u].x_off);
        const long long b = std::min(end, p.utts[u].x_off + p.utts[u].length);
        pieces.push_back(CopyPiece{hx + a, xs[u] + (a - p.utts[u].x_off), b - a});
        if (b < p.utts[u].x_off + p.utts[u].length) break;  // continues in the next slice
      }
      finite &= copy_finite_parallel(pieces.data(), (int)pieces.size());
      check(hipMemcpyAsync(d + at, hx + at, (end - at) * sizeof(double), hipMemcpyHostToDevice, s),
            "upload x");
    }
    if (!finite) {
      // No kernel has been queued; the uploads finish before hx is reused.
      check(hipStreamSynchronize(s), "upload x");
      invalid("x must contain only finite samples");
    }
    dx = d;
  }
  *staged = dx;
  return p;
}

// stage_batch, then decimation and mean removal launch by launch: on return
// the decimated, mean-removed y is on the device (frozen-decision entry).
Plan preprocess_batch(hv_workspace* ws, int count, const double* const* xs,
                      const int64_t* lengths, int fs, double f0_floor, double f0_ceil,
                      double* timings, bool device_x) {
  const double* dx = nullptr;
  Plan p = stage_batch(ws, count, xs, lengths, fs, f0_floor, f0_ceil, timings, device_x, &dx);
  enqueue_preprocess(ws, p, size_buffers(ws, p, dx), true);
  return p;
}

void run_batch(hv_workspace* ws, int count, const double* const* xs, const int64_t* lengths,
               int fs, double f0_floor, double f0_ceil, int32_t* offsets, double* best,
               int64_t* nnz, double* timings, bool device_x, bool defer = false) {
  if (!ws) invalid("null workspace");
  if (ws->submitted) invalid("the workspace has a submitted analysis to collect first");
  DeviceGuard guard(ws->device);
  const double* dx = nullptr;
  const Plan p = stage_batch(ws, count, xs, lengths, fs, f0_floor, f0_ceil, timings, device_x,
                             &dx);
  analyze_device(ws, p, dx, offsets, best, nnz, timings, defer);
}

void submit_batch(hv_workspace* ws, int count, const double* const* xs, const int64_t* lengths,
                  int fs, double f0_floor, double f0_ceil, bool timings, bool device_x) {
  if (!ws) invalid("null workspace");
  if (ws->submitted) invalid("the workspace has a submitted analysis to collect first");
  DeviceGuard guard(ws->device);
  double on = 0.0;
  const double* dx = nullptr;
  const Plan p = stage_batch(ws, count, xs, lengths, fs, f0_floor, f0_ceil,
                             timings ? &on : nullptr, device_x, &dx);
  if (ws->after_pending) {
    check(hipStreamWaitEvent(ws->stream, ws->after_ev, 0), "hipStreamWaitEvent");
    ws->after_pending = false;
  }
  analyze_submit(ws, p, dx, true);
}

// Records the buffer sizes of each group's analysis (two-step form), so that
// the next submission grows every buffer once to the largest of them.
void reserve_batches(hv_workspace* ws, int groups, const int* counts, const int64_t* lengths,
                     int fs, double f0_floor, double f0_ceil, bool host_input) {
  if (!ws) invalid("null workspace");
  if (groups < 1) invalid("groups must be positive");
  DeviceGuard guard(ws->device);
  long long at = 0;
  for (int g = 0; g < groups; ++g) {
    if (counts[g] < 1) invalid("count must be positive");
    const Plan p = plan_batch(ws, counts[g], lengths + at, fs, f0_floor, f0_ceil);
    at += counts[g];
    Recording rec;
    ws->h_meta.ensure<char>(plan_bytes(p.cfg.count));
    ws->plan.ensure<char>(plan_bytes(p.cfg.count));
    if (host_input) {
      ws->h_x.ensure<double>(p.total_x);
      ws->x.ensure<double>(p.total_x);
    }
    Bufs b = size_buffers(ws, p, nullptr);
    size_candidate_copy(ws, p, b);
  }
}

// Frozen-decision refinement of entries (value and/or adjoint), then the
// adjoint of mean removal and decimation.
void refine_entries(hv_workspace* ws, int count, const double* const* xs,
                    const int64_t* lengths, int fs, double f0_floor, double f0_ceil,
                    int64_t n_entries, const int64_t* frame, const double* f0_initial,
                    const double* weight, double* refined, double* grad_x) {
  if (!ws) invalid("null workspace");
  if (ws->submitted) invalid("the workspace has a submitted analysis to collect first");
  DeviceGuard guard(ws->device);
  if (n_entries < 0) invalid("entry count must be nonnegative");
  if (n_entries > 0 && (!frame || !f0_initial)) invalid("null entry arrays");
  if (grad_x && n_entries > 0 && !weight) invalid("grad_x needs weights");
  const Plan p = preprocess_batch(ws, count, xs, lengths, fs, f0_floor, f0_ceil, nullptr, false);
  const Config& c = p.cfg;
  hipStream_t s = ws->stream;
  const Utt* dutt = ws->d_utt;
  for (int64_t e = 0; e < n_entries; ++e) {
    if (frame[e] < 0 || frame[e] >= p.total_frames)
      invalid("entry frame " + std::to_string(frame[e]) + " is outside the batch's frames");
    if (!(f0_initial[e] >= c.f0_floor * (1.0 - 1e-9) && f0_initial[e] <= c.f0_ceil * (1.0 + 1e-9)))
      invalid("entry initial f0 outside [f0_floor, f0_ceil]");
  }
  // Entry arrays through one pinned staging buffer: frames, f0, weights.
  const size_t ne = (size_t)n_entries;
  char* h = ws->h_ent.ensure<char>(std::max<size_t>(1, ne * (sizeof(long long) + 2 * sizeof(double))));
  long long* hf = reinterpret_cast<long long*>(h);
  double* hi = reinterpret_cast<double*>(h + ne * sizeof(long long));
  double* hw = hi + ne;
  for (size_t e = 0; e < ne; ++e) {
    hf[e] = frame[e];
    hi[e] = f0_initial[e];
    hw[e] = weight ? weight[e] : 0.0;
  }
  long long* dframe = ws->ent_frame.ensure<long long>(std::max<size_t>(1, ne));
  double* dinit = ws->ent_f0.ensure<double>(std::max<size_t>(1, ne));
  double* dw = ws->ent_w.ensure<double>(std::max<size_t>(1, ne));
  double* dout = ws->ent_out.ensure<double>(std::max<size_t>(1, ne));
  if (ne > 0) {
    check(hipMemcpyAsync(dframe, hf, ne * sizeof(long long), hipMemcpyHostToDevice, s), "upload entries");
    check(hipMemcpyAsync(dinit, hi, ne * sizeof(double), hipMemcpyHostToDevice, s), "upload entries");
    check(hipMemcpyAsync(dw, hw, ne * sizeof(double), hipMemcpyHostToDevice, s), "upload entries");
  }
  double* gy = nullptr;
  if (grad_x) {
    gy = ws->gy.ensure<double>(p.total_y);
    launch_fill_int(s, reinterpret_cast<int*>(gy), 2 * p.total_y, 0);
  }
  launch_refine_entries(s, c, dutt, ws->d_f_off, n_entries, dframe, dinit,
                        grad_x ? dw : nullptr, ws->y.get<double>(), refined ? dout : nullptr, gy);
  if (refined && ne > 0) {
    check(hipMemcpyAsync(hw, dout, ne * sizeof(double), hipMemcpyDeviceToHost, s), "download refined");
  }
  if (grad_x) {
    // Mean removal is symmetric: y_bar minus its mean.
    launch_remove_mean(s, c, dutt, ws->d_p_off, p.total_pblocks,
                       ws->d_y_off, p.total_y, gy,
                       ws->partial.ensure<double>(p.total_pblocks), ws->mean.ensure<double>(count));
    double* gx = ws->gx.ensure<double>(p.total_x);
    if (c.ratio > 1) {
      launch_decimate_transpose(
          s, c, dutt, ws->d_c_off, p.total_chunks, p.max_chunks, ws->d_x_off, p.total_x, gy,
          ws->pass.ensure<double>(p.total_m), ws->states.ensure<double>(3 * p.total_chunks),
          ws->carries.ensure<double>(3 * p.total_chunks),
          ws->carry_tmp.ensure<double>(carry_scratch_doubles(count, p.max_chunks)),
          ws->full.ensure<double>(p.total_m), gx);
    } else {
      check(hipMemcpyAsync(gx, gy, p.total_x * sizeof(double), hipMemcpyDeviceToDevice, s),
            "copy gradient");
    }
    double* hgx = ws->h_gx.ensure<double>(p.total_x);
    check(hipMemcpyAsync(hgx, gx, p.total_x * sizeof(double), hipMemcpyDeviceToHost, s),
          "download gradient");
    check(hipStreamSynchronize(s), "adjoint");
    std::memcpy(grad_x, hgx, p.total_x * sizeof(double));
  } else {
    check(hipStreamSynchronize(s), "refine entries");
  }
  if (refined) std::memcpy(refined, hw, ne * sizeof(double));
  // The analysis result of this workspace is no longer current.
  ws->last_nnz = -1;
}

}  // namespace
}  // namespace hv

extern "C" {

const char* hv_last_error(void) { return g_error.c_str(); }

int hv_device_supported(int device) { return Device::supported(device); }

hv_status hv_workspace_create(int device, hv_workspace** out) {
  return guarded([&] {
    if (!out) invalid("null output pointer");
    *out = nullptr;
    Device::validate(device);
    DeviceGuard guard(device);
    *out = new hv_workspace(device);
  });
}

void hv_workspace_destroy(hv_workspace* ws) {
  if (!ws) return;
  try {
    DeviceGuard guard(ws->device);
    delete ws;
  } catch (...) {
    // Best effort after driver shutdown.
  }
}

int hv_workspace_device(const hv_workspace* ws) { return ws ? ws->device : -1; }

hv_status hv_decimate(hv_workspace* ws, const double* x, int64_t length, int ratio, double* y,
                      int64_t y_length) {
  return guarded([&] {
    if (!ws || !x || !y) invalid("null pointer");
    if (ws->submitted) invalid("the workspace has a submitted analysis to collect first");
    if (ratio < 1 || ratio > 12) invalid("ratio must be between 1 and 12");
    if (length < 2LL * ratio || length > 0x7fffffffLL) invalid("invalid length");
    if (y_length != (length + ratio - 1) / ratio) invalid("y_length must be ceil(length / ratio)");
    DeviceGuard guard(ws->device);
    Plan p;
    fill_config(p.cfg, 1, ratio, 8000 * ratio, 8000.0, 71.0, 800.0, 0, 0, 0);
    layout(p, {length}, {(int)y_length}, {0}, {0});
    hipStream_t s = ws->stream;
    upload_plan(ws, p);
    double* hx = ws->h_x.ensure<double>(length);
    std::memcpy(hx, x, length * sizeof(double));
    double* dx = ws->x.ensure<double>(length);
    check(hipMemcpyAsync(dx, hx, length * sizeof(double), hipMemcpyHostToDevice, s), "upload x");
    double* dy = ws->y.ensure<double>(y_length);
    if (ratio > 1)
      launch_decimate(s, p.cfg, ws->d_utt, ws->d_c_off, p.total_chunks, p.max_chunks, dx,
                      ws->pass.ensure<double>(p.total_m), ws->states.ensure<double>(3 * p.total_chunks),
                      ws->carries.ensure<double>(3 * p.total_chunks),
                      ws->carry_tmp.ensure<double>(carry_scratch_doubles(1, p.max_chunks)), dy);
    else
      launch_copy_input(s, p.cfg, ws->d_utt, ws->d_y_off, p.total_y, dx, dy);
    double* hy = ws->h_best.ensure<double>(y_length);
    check(hipMemcpyAsync(hy, dy, y_length * sizeof(double), hipMemcpyDeviceToHost, s), "download y");
    check(hipStreamSynchronize(s), "decimate");
    std::memcpy(y, hy, y_length * sizeof(double));
  });
}

hv_status hv_decimate_device_noalloc(int device, const double* x_device, int64_t length,
                                     int ratio, double* y_device, int64_t y_length,
                                     void* stream) {
  return guarded([&] {
    if (!x_device || !y_device) invalid("null pointer");
    if (ratio < 1 || ratio > 12) invalid("ratio must be between 1 and 12");
    if (length < 2LL * ratio || length > 0x7fffffffLL) invalid("invalid length");
    if (y_length != (length + ratio - 1) / ratio) invalid("y_length must be ceil(length / ratio)");
    Device::validate(device);
    int devices = 0;
    check(hipGetDeviceCount(&devices), "hipGetDeviceCount");
    // The plan of one utterance, as layout() builds it for a batch of one.
    Config cfg{};
    fill_config(cfg, 1, ratio, 8000 * ratio, 8000.0, 71.0, 800.0, 0, 0, 0);
    Utt u{};
    u.length = length;
    u.n = (int)y_length;
    if (ratio > 1) {
      const long long padded = length + 2LL * cfg.lag;
      u.m = padded + 18;
      const long long nout = (padded - 1) / ratio + 1;
      u.nbeg = (int)(ratio - ratio * nout + padded);
    }
    u.chunks = (int)((u.m + kChunk - 1) / kChunk);
    // Scratch layout: plan words, pass buffer, chunk states, carries, carry
    // scratch; each part 256-byte aligned.
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    constexpr int kUttWords = (int)(sizeof(Utt) / 8);
    static_assert(sizeof(Utt) % 8 == 0 && kUttWords + 4 <= kStoreWords, "plan words");
    const size_t at_pass = 256;
    const size_t at_states = at_pass + up((size_t)u.m * sizeof(double));
    const size_t at_carries = at_states + up((size_t)3 * u.chunks * sizeof(double));
    const size_t at_tmp = at_carries + up((size_t)3 * u.chunks * sizeof(double));
    const size_t need = at_tmp + up(carry_scratch_doubles(1, u.chunks) * sizeof(double));
    if (need > kScratchBytes)
      invalid("input needs " + std::to_string(need) + " bytes of scratch; the allocation-free "
              "entry point holds " + std::to_string(kScratchBytes) +
              " (inputs up to about 7.6 million samples): use hv_decimate");
    DeviceGuard guard(device);
    char* base = static_cast<char*>(scratch_base(device, devices));
    hipStream_t s = static_cast<hipStream_t>(stream);
    // The utterance, then its chunk and output offset tables ([0, total]).
    StoreWords words{};
    std::memcpy(words.w, &u, sizeof(Utt));
    words.w[kUttWords + 0] = 0;
    words.w[kUttWords + 1] = (unsigned long long)u.chunks;
    words.w[kUttWords + 2] = 0;
    words.w[kUttWords + 3] = (unsigned long long)y_length;
    auto* plan = reinterpret_cast<unsigned long long*>(base);
    launch_store_words(s, words, kUttWords + 4, plan);
    const Utt* d_utt = reinterpret_cast<const Utt*>(plan);
    const long long* d_c_off = reinterpret_cast<const long long*>(plan + kUttWords);
    const long long* d_y_off = reinterpret_cast<const long long*>(plan + kUttWords + 2);
    if (ratio > 1)
      launch_decimate(s, cfg, d_utt, d_c_off, u.chunks, u.chunks, x_device,
                      reinterpret_cast<double*>(base + at_pass),
                      reinterpret_cast<double*>(base + at_states),
                      reinterpret_cast<double*>(base + at_carries),
                      reinterpret_cast<double*>(base + at_tmp), y_device);
    else
      launch_copy_input(s, cfg, d_utt, d_y_off, y_length, x_device, y_device);
  });
}

hv_status hv_analyze(hv_workspace* ws, const double* y, int64_t y_length, int fft_size,
                     double fs_work, const double* boundaries, int bands, int frames,
                     double f0_floor, double f0_ceil, int capacity, int32_t* offsets,
                     double* best, int64_t* nnz, double* timings_ms) {
  return guarded([&] {
    if (!ws || !y || !boundaries || !nnz) invalid("null pointer");
    if (ws->submitted) invalid("the workspace has a submitted analysis to collect first");
    if (y_length < 2 || y_length > 0x7fffffffLL || bands < 3 || frames < 1 || capacity < 1)
      invalid("invalid analysis sizes");
    if (fft_size < y_length || (fft_size & (fft_size - 1)) != 0)
      invalid("fft_size must be a power of two no smaller than y_length");
    if (!std::isfinite(fs_work) || fs_work <= 0) invalid("fs_work must be finite and positive");
    if (!std::isfinite(f0_floor) || !std::isfinite(f0_ceil) || f0_floor <= 0 || f0_ceil <= f0_floor)
      invalid("require finite 0 < f0_floor < f0_ceil");
    if (f0_ceil >= fs_work / 2) invalid("f0_ceil must be below the internal Nyquist frequency");
    for (int b = 0; b < bands; ++b)
      if (!std::isfinite(boundaries[b]) || boundaries[b] <= 0)
        invalid("band boundaries must be finite and positive");
    // Frames map to filterbank tiles with integer arithmetic: fs_work = num / den.
    int den = 0;
    for (int d = 1; d <= 12 && !den; ++d)
      if (std::floor(fs_work * d) == fs_work * d) den = d;
    if (!den) invalid("fs_work must be an integer sample rate divided by 1..12");
    DeviceGuard guard(ws->device);
    prepare_filters(ws, fs_work, f0_floor, f0_ceil, boundaries, bands);
    prepare_twiddles(ws, fs_work, f0_floor);
    prepare_bintable(ws, fs_work, f0_floor, f0_ceil);
    Plan p;
    fill_config(p.cfg, 1, den, (int)(fs_work * den), fs_work, f0_floor, f0_ceil, bands, capacity,
                ws->twiddle_n);
    apply_bintable(ws, p.cfg);
    p.cfg.lds_limit = ws->lds_limit;
    p.cfg.lag = 0;
    layout(p, {y_length}, {(int)y_length}, {fft_size}, {frames});
    prepare_phase(ws, p);
    hipStream_t s = ws->stream;
    start_timers(ws, timings_ms != nullptr);
    check(hipEventRecord(ws->ev[0], s), "hipEventRecord");
    upload_plan(ws, p);
    double* hy = ws->h_x.ensure<double>(y_length);
    std::memcpy(hy, y, y_length * sizeof(double));
    check(hipMemcpyAsync(ws->y.ensure<double>(y_length), hy, y_length * sizeof(double),
                         hipMemcpyHostToDevice, s), "upload y");
    check(hipEventRecord(ws->ev[1], s), "hipEventRecord");
    analyze_device(ws, p, nullptr, offsets, best, nnz, timings_ms);
  });
}

hv_status hv_analyze_batch(hv_workspace* ws, int count, const double* const* xs,
                           const int64_t* lengths, int fs, double f0_floor, double f0_ceil,
                           int32_t* offsets, double* best, int64_t* nnz, double* timings_ms) {
  return guarded([&] {
    if (!xs || !lengths || !nnz) invalid("null pointer");
    for (int u = 0; u < count; ++u)
      if (!xs[u]) invalid("null waveform pointer");
    run_batch(ws, count, xs, lengths, fs, f0_floor, f0_ceil, offsets, best, nnz, timings_ms, false);
  });
}

hv_status hv_analyze_batch_device(hv_workspace* ws, int count, const double* x_device,
                                  const int64_t* lengths, int fs, double f0_floor,
                                  double f0_ceil, int32_t* offsets, double* best, int64_t* nnz,
                                  double* timings_ms) {
  return guarded([&] {
    if (!x_device || !lengths || !nnz) invalid("null pointer");
    const double* xs[1] = {x_device};
    run_batch(ws, count, xs, lengths, fs, f0_floor, f0_ceil, offsets, best, nnz, timings_ms, true);
  });
}

hv_status hv_analyze_batch_async(hv_workspace* ws, int count, const double* const* xs,
                                 const int64_t* lengths, int fs, double f0_floor,
                                 double f0_ceil, int32_t* offsets, double* best, int64_t* nnz,
                                 double* timings_ms) {
  return guarded([&] {
    if (!xs || !lengths || !nnz) invalid("null pointer");
    for (int u = 0; u < count; ++u)
      if (!xs[u]) invalid("null waveform pointer");
    run_batch(ws, count, xs, lengths, fs, f0_floor, f0_ceil, offsets, best, nnz, timings_ms, false,
              true);
  });
}

hv_status hv_analyze_batch_device_async(hv_workspace* ws, int count, const double* x_device,
                                        const int64_t* lengths, int fs, double f0_floor,
                                        double f0_ceil, int32_t* offsets, double* best,
                                        int64_t* nnz, double* timings_ms) {
  return guarded([&] {
    if (!x_device || !lengths || !nnz) invalid("null pointer");
    const double* xs[1] = {x_device};
    run_batch(ws, count, xs, lengths, fs, f0_floor, f0_ceil, offsets, best, nnz, timings_ms, true,
              true);
  });
}

hv_status hv_analyze_batch_submit(hv_workspace* ws, int count, const double* const* xs,
                                  const int64_t* lengths, int fs, double f0_floor,
                                  double f0_ceil, int timings) {
  return guarded([&] {
    if (!xs || !lengths) invalid("null pointer");
    for (int u = 0; u < count; ++u)
      if (!xs[u]) invalid("null waveform pointer");
    submit_batch(ws, count, xs, lengths, fs, f0_floor, f0_ceil, timings != 0, false);
  });
}

hv_status hv_analyze_batch_device_submit(hv_workspace* ws, int count, const double* x_device,
                                         const int64_t* lengths, int fs, double f0_floor,
                                         double f0_ceil, int timings) {
  return guarded([&] {
    if (!x_device || !lengths) invalid("null pointer");
    const double* xs[1] = {x_device};
    submit_batch(ws, count, xs, lengths, fs, f0_floor, f0_ceil, timings != 0, true);
  });
}

hv_status hv_reserve_batches(hv_workspace* ws, int groups, const int* counts,
                             const int64_t* lengths, int fs, double f0_floor, double f0_ceil,
                             int host_input) {
  return guarded([&] {
    if (!counts || !lengths) invalid("null pointer");
    if (ws && ws->submitted) invalid("the workspace has a submitted analysis to collect first");
    reserve_batches(ws, groups, counts, lengths, fs, f0_floor, f0_ceil, host_input != 0);
  });
}

hv_status hv_analyze_batch_collect(hv_workspace* ws, int32_t* offsets, double* best,
                                   int64_t* nnz, double* timings_ms) {
  return guarded([&] {
    if (!ws || !nnz) invalid("null pointer");
    if (timings_ms && !ws->timing) invalid("the analysis was submitted without timings");
    DeviceGuard guard(ws->device);
    analyze_collect(ws, offsets, best, nnz, timings_ms, true);
  });
}

hv_status hv_wait_candidates(hv_workspace* ws, int part) {
  return guarded([&] {
    if (!ws) invalid("null workspace");
    if (part != 0 && part != 1) invalid("part must be 0 (f0) or 1 (f0 and scores)");
    check(hipEventSynchronize(ws->cand_ev[part]), "wait for candidates");
  });
}

hv_status hv_fetch_initial(hv_workspace* ws, double* f0_initial) {
  return guarded([&] {
    if (!ws) invalid("null workspace");
    if (ws->last_nnz < 0) invalid("no analysis result to fetch");
    if (ws->last_nnz == 0) return;
    if (!f0_initial) invalid("null pointer");
    DeviceGuard guard(ws->device);
    check(hipMemcpyAsync(f0_initial, ws->cinit.get<double>(), ws->last_nnz * sizeof(double),
                         hipMemcpyDeviceToHost, ws->stream), "download initial f0");
    check(hipStreamSynchronize(ws->stream), "download initial f0");
  });
}

hv_status hv_refine_entries(hv_workspace* ws, int count, const double* const* xs,
                            const int64_t* lengths, int fs, double f0_floor, double f0_ceil,
                            int64_t n_entries, const int64_t* frame, const double* f0_initial,
                            const double* weight, double* refined, double* grad_x) {
  return guarded([&] {
    if (!xs || !lengths) invalid("null input arrays");
    for (int u = 0; u < count; ++u)
      if (!xs[u]) invalid("null utterance pointer");
    refine_entries(ws, count, xs, lengths, fs, f0_floor, f0_ceil, n_entries, frame, f0_initial,
                   weight, refined, grad_x);
  });
}

hv_status hv_get_tuning(hv_workspace* ws, hv_tuning* out) {
  return guarded([&] {
    if (!ws || !out) invalid("null pointer");
    *out = ws->tune;
  });
}

hv_status hv_set_tuning(hv_workspace* ws, const hv_tuning* in) {
  return guarded([&] {
    if (!ws || !in) invalid("null pointer");
    if (ws->submitted) invalid("the workspace has a submitted analysis to collect first");
    if (in->fft_min_half < 1) invalid("fft_min_half must be positive");
    if (in->graphs != 0 && in->graphs != 1) invalid("graphs must be 0 or 1");
    if (in->slab_budget_mib < 64) invalid("slab_budget_mib must be at least 64");
    if (in->refine_table != 0 && in->refine_table != 1) invalid("refine_table must be 0 or 1");
    if (in->refine_groups != 0 && (in->refine_groups < 64 || in->refine_groups > 4096))
      invalid("refine_groups must be 0 or between 64 and 4096");
    ws->tune = *in;  // a new band split takes effect at the next call (filter cache key)
  });
}

hv_status hv_order_after(hv_workspace* ws, hv_workspace* before) {
  return guarded([&] {
    if (!ws || !before) invalid("null workspace");
    if (ws == before) invalid("a workspace is ordered after itself already");
    if (ws->device != before->device) invalid("workspaces on different devices");
    DeviceGuard guard(ws->device);
    check(hipEventRecord(ws->after_ev, before->stream), "hipEventRecord");
    ws->after_pending = true;
  });
}

void hv_allocation_stats(int64_t* allocations, int64_t* device_regrowths) {
  if (allocations) *allocations = g_allocations.load(std::memory_order_relaxed);
  if (device_regrowths) *device_regrowths = g_device_regrowths.load(std::memory_order_relaxed);
}

hv_status hv_graph_stats(hv_workspace* ws, int64_t* replays, int64_t* captures) {
  return guarded([&] {
    if (!ws || !replays || !captures) invalid("null pointer");
    *replays = ws->graph_hits;
    *captures = ws->graph_captures;
  });
}

hv_status hv_last_stats(hv_workspace* ws, int64_t* stats) {
  return guarded([&] {
    if (!ws || !stats) invalid("null pointer");
    if (ws->last_nnz < 0) invalid("no analysis result");
    const unsigned long long* h = ws->h_work.get<unsigned long long>();
    for (int k = 0; k < HV_STAT_SLOTS; ++k) stats[k] = h ? (int64_t)h[k] : 0;
  });
}

hv_status hv_fetch_candidates(hv_workspace* ws, double* f0, double* scores) {
  return guarded([&] {
    if (!ws) invalid("null workspace");
    if (ws->last_nnz < 0) invalid("no analysis result to fetch");
    if (ws->last_nnz == 0) return;
    if (!f0 || !scores) invalid("null pointer");
    check(hipEventSynchronize(ws->cand_ev[1]), "wait for candidates");
    const long long n = ws->last_nnz;
    const double* h = ws->h_cand.get<double>();
    std::memcpy(f0, h, n * sizeof(double));
    std::memcpy(scores, h + ws->cand_capacity, n * sizeof(double));
  });
}

hv_status hv_result_view(hv_workspace* ws, const int32_t** offsets, const double** best,
                         const double** f0, const double** scores, int64_t* nnz) {
  return guarded([&] {
    if (!ws || !offsets || !best || !f0 || !scores || !nnz) invalid("null pointer");
    if (ws->last_nnz < 0) invalid("no analysis result to view");
    *offsets = ws->h_offsets.get<int32_t>();
    *best = ws->h_best.get<double>();
    *f0 = ws->h_cand.get<double>();
    *scores = ws->h_cand.get<double>() + ws->cand_capacity;
    *nnz = ws->last_nnz;
  });
}

}  // extern "C"
