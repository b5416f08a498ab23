// Device selection, guard and error mapping of the backend (device_check.h)
// driven by a fake HIP runtime on the CPU.  The cases follow the reference's
// tests/cpp/test_cuda_device.cpp:9-101, restated for HIP's model (the state
// a call restores is the caller's current device, not a context).
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "device_check.h"

namespace {

struct Fake {
  using Error = int;
  static constexpr Error kSuccess = 0, kOutOfMemory = 2, kNoDevice = 100, kInsufficientDriver = 35;
  static int count;
  static Error count_status;
  static std::vector<std::string> arch;
  static int last_queried;
  static int current;
  static std::vector<int> switches;
  static bool names;
  static Error fail_get;
  static int fail_set_to;  // set_device(fail_set_to) fails
  static int sticky;       // last error not yet cleared
  static Error device_count(int* n) {
    *n = count;
    return count_status;
  }
  static Error properties(int device, std::string* a, std::string* name) {
    last_queried = device;
    *a = arch[device];
    *name = "test GPU";
    return kSuccess;
  }
  static Error get_device(int* d) {
    if (fail_get) {
      sticky = fail_get;
      return fail_get;
    }
    *d = current;
    return kSuccess;
  }
  static Error set_device(int d) {
    if (d == fail_set_to) {
      sticky = 101;
      return 101;
    }
    switches.push_back(d);
    current = d;
    return kSuccess;
  }
  static const char* error_name(Error e) {
    if (!names) return nullptr;
    switch (e) {
      case kOutOfMemory: return "hipErrorOutOfMemory";
      case kNoDevice: return "hipErrorNoDevice";
      case kInsufficientDriver: return "hipErrorInsufficientDriver";
      default: return "hipErrorInvalidDevice";
    }
  }
  static const char* error_string(Error) { return names ? "test description" : nullptr; }
  static void clear_error() { sticky = 0; }
};
int Fake::count = 1;
Fake::Error Fake::count_status = 0;
std::vector<std::string> Fake::arch = {"gfx950:sramecc+:xnack-", "gfx950", "gfx950"};
int Fake::last_queried = -1;
int Fake::current = 0;
std::vector<int> Fake::switches;
bool Fake::names = true;
Fake::Error Fake::fail_get = 0;
int Fake::fail_set_to = -1;
int Fake::sticky = 0;

using Ops = hv::DeviceOps<Fake>;
int failures = 0;

void require(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++failures;
  }
}

// The operation throws hv::Failure with `status` and a message containing `text`.
template <class F> void fails_with(F&& op, hv_status status, const char* text, const char* what) {
  try {
    op();
  } catch (const hv::Failure& f) {
    const bool ok = f.status == status && std::string(f.what()).find(text) != std::string::npos;
    if (!ok) std::fprintf(stderr, "  got status %d: %s\n", (int)f.status, f.what());
    require(ok, what);
    return;
  }
  require(false, what);
}

}  // namespace

int main() {
  // Availability follows visible indices and the gfx950 architecture.
  Ops::validate(0);
  require(Ops::supported(0) == 1, "gfx950 with feature suffixes is supported");
  Fake::arch[0] = "gfx942";
  fails_with([] { Ops::validate(0); }, HV_EUNSUPPORTED, "(test GPU) is gfx942", "unsupported arch");
  require(Ops::supported(0) == 0, "unsupported arch is not available");
  Fake::arch[0] = "gfx950";
  Fake::count = 0;
  fails_with([] { Ops::validate(0); }, HV_ENODEV, "HIP_VISIBLE_DEVICES", "no visible GPU");
  require(Ops::supported(0) == 0, "no GPU: not available");
  Fake::count = 1;
  Fake::count_status = Fake::kInsufficientDriver;
  fails_with([] { Ops::validate(0); }, HV_ENODEV, "Install a ROCm driver", "insufficient driver");
  Fake::names = false;  // a runtime without descriptions still names the code
  fails_with([] { Ops::validate(0); }, HV_ENODEV, "HIP: 35", "numeric code without names");
  fails_with([] { Ops::validate(0); }, HV_ENODEV, "Install a ROCm driver", "driver hint without names");
  Fake::names = true;
  Fake::sticky = 9;
  require(Ops::supported(0) == 0 && Fake::sticky == 0, "supported() never fails and clears the error");
  Fake::count_status = Fake::kSuccess;
  fails_with([] { Ops::check(Fake::kOutOfMemory, "hipMalloc"); }, HV_ENOMEM,
             "GPU memory allocation failed", "out of memory");
  fails_with([] { Ops::check(7, "hipMemcpyAsync"); }, HV_EDEVICE, "(hipMemcpyAsync)", "other errors");
  fails_with([] { Ops::validate(-1); }, HV_EINVAL, "nonnegative", "negative index");
  fails_with([] { Ops::validate(1); }, HV_EINVAL, "out of range", "index beyond the visible GPUs");
  require(Ops::supported(-1) == 0 && Ops::supported(1) == 0, "invalid indices are not available");
  Fake::count = 2;
  Fake::arch[0] = "gfx90a";
  Ops::validate(1);
  require(Fake::last_queried == 1, "the requested device is the one queried");
  fails_with([] { Ops::validate(0); }, HV_EUNSUPPORTED, "device 0", "the message names the device");
  Fake::arch[0] = "gfx950";

  // The guard switches to the device and restores the caller's.
  Fake::count = 3;
  Fake::current = 2;
  Fake::switches.clear();
  {
    Ops::Guard g(0);
    require(Fake::current == 0, "guard makes the device current");
    {
      Ops::Guard nested(1);
      require(Fake::current == 1, "nested guard");
    }
    require(Fake::current == 0, "nested guard restores");
  }
  require(Fake::current == 2, "guard restores the caller's device");
  require(Fake::switches == std::vector<int>({0, 1, 0, 2}), "switch sequence");
  Fake::switches.clear();
  try {
    Ops::Guard g(0);
    throw std::runtime_error("processing failure");
  } catch (const std::runtime_error&) {
  }
  require(Fake::current == 2 && Fake::switches == std::vector<int>({0, 2}),
          "restored when the guarded work throws");
  Fake::switches.clear();
  { Ops::Guard g(2); }
  require(Fake::switches.empty(), "no switch when the device is already current");
  Fake::fail_get = 11;
  fails_with([] { Ops::Guard g(0); }, HV_EDEVICE, "hipGetDevice", "failed query throws");
  require(Fake::current == 2 && Fake::switches.empty() && Fake::sticky == 0,
          "a failed query changes nothing and leaves no sticky error");
  Fake::fail_get = 0;
  Fake::fail_set_to = 0;
  fails_with([] { Ops::Guard g(0); }, HV_EDEVICE, "hipSetDevice", "failed switch throws");
  require(Fake::current == 2 && Fake::sticky == 0, "a failed switch keeps the caller's device");
  Fake::fail_set_to = 2;  // restoring fails: the destructor swallows it
  {
    Ops::Guard g(1);
  }
  require(Fake::current == 1 && Fake::sticky == 0, "a failed restore is dropped, not thrown");
  Fake::fail_set_to = -1;
  if (failures) return 1;
  std::puts("device tests passed");
  return 0;
}
