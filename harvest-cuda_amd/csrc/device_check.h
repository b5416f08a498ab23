// Device selection, device guard and error mapping of the gfx950 backend
// (counterpart of the reference's src/cuda_device.h:8-108: check_cuda,
// check_harvest_device and CudaContextGuard).  HIP has no primary-context
// handle to retain: the state a call must restore is the caller's current
// device.  Everything here is written against a runtime policy R so that the
// CPU tests (tests/cpp/test_device.cpp) can drive it with a fake runtime;
// the backend instantiates it with the HIP runtime (capi.hip).
//
// R provides:
//   using Error = ...;                     the runtime's error code type
//   static constexpr Error kSuccess, kOutOfMemory, kNoDevice, kInsufficientDriver;
//   static Error device_count(int*);
//   static Error properties(int device, std::string* arch, std::string* name);
//   static Error get_device(int*);
//   static Error set_device(int);
//   static const char* error_name(Error);   may return nullptr
//   static const char* error_string(Error); may return nullptr
//   static void clear_error();              drops the sticky last error
#pragma once
#include <stdexcept>
#include <string>

#include "../../include/harvest_hip.h"

namespace hv {

// A failure with the status the C ABI returns for it.
struct Failure : std::runtime_error {
  hv_status status;
  Failure(hv_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] inline void invalid(const std::string& m) { throw Failure(HV_EINVAL, m); }

template <class R> struct DeviceOps {
  using Error = typename R::Error;

  // Throws the status and message of a failed runtime call.
  static void check(Error e, const char* what) {
    if (e == R::kSuccess) return;
    R::clear_error();
    hv_status st = HV_EDEVICE;
    if (e == R::kOutOfMemory) st = HV_ENOMEM;
    if (e == R::kNoDevice || e == R::kInsufficientDriver) st = HV_ENODEV;
    // Runtimes without a driver can return no descriptions.
    const char* name = R::error_name(e);
    const char* detail = R::error_string(e);
    std::string m = std::string("HIP: ") + (name ? name : std::to_string((int)e));
    if (detail) m += std::string(": ") + detail;
    m += std::string(" (") + what + ")";
    if (e == R::kInsufficientDriver)
      m += ". Install a ROCm driver that supports this HIP runtime.";
    else if (st == HV_ENOMEM)
      m += ". GPU memory allocation failed; try shorter audio or release unused GPU allocations.";
    else if (st == HV_ENODEV)
      m += ". Check the ROCm driver and HIP_VISIBLE_DEVICES before starting Python.";
    throw Failure(st, m);
  }

  // 1 if `device` is a visible gfx950 GPU, else 0; never throws and leaves
  // no sticky error (cuda_available, src/cuda_backend.cu:794-803).
  static int supported(int device) noexcept {
    try {
      if (device < 0) return 0;
      int n = 0;
      if (R::device_count(&n) != R::kSuccess) {
        R::clear_error();
        return 0;
      }
      if (device >= n) return 0;
      std::string arch, name;
      if (R::properties(device, &arch, &name) != R::kSuccess) {
        R::clear_error();
        return 0;
      }
      return arch.compare(0, 6, "gfx950") == 0 ? 1 : 0;
    } catch (...) {
      return 0;
    }
  }

  // The checks of workspace creation (check_harvest_device,
  // src/cuda_device.h:26-40): indices follow HIP_VISIBLE_DEVICES.
  static void validate(int device) {
    if (device < 0) invalid("device must be a nonnegative visible GPU index");
    int n = 0;
    check(R::device_count(&n), "hipGetDeviceCount");
    if (n == 0) check(R::kNoDevice, "no visible GPU");
    if (device >= n)
      invalid("device " + std::to_string(device) + " is out of range; " + std::to_string(n) +
              " GPU(s) are visible through HIP_VISIBLE_DEVICES");
    std::string arch, name;
    check(R::properties(device, &arch, &name), "hipGetDeviceProperties");
    if (arch.compare(0, 6, "gfx950") != 0)
      throw Failure(HV_EUNSUPPORTED, "Harvest requires an MI355X-class gfx950 GPU; device " +
                                         std::to_string(device) + " (" + name + ") is " + arch +
                                         ". Select a supported GPU with HIP_VISIBLE_DEVICES "
                                         "before starting Python.");
  }

  // Makes `device` current for the guard's scope and restores the caller's
  // device on every exit path (CudaContextGuard, src/cuda_device.h:92-108).
  // No switch happens when the device is already current; a failed query
  // throws before anything changes; a failed switch throws with the
  // caller's device still current.
  class Guard {
   public:
    explicit Guard(int device) {
      check(R::get_device(&previous_), "hipGetDevice");
      if (previous_ != device) {
        check(R::set_device(device), "hipSetDevice");
        switched_ = true;
      }
    }
    Guard(const Guard&) = delete;
    Guard& operator=(const Guard&) = delete;
    ~Guard() {
      if (switched_ && R::set_device(previous_) != R::kSuccess) R::clear_error();
    }

   private:
    int previous_ = -1;
    bool switched_ = false;
  };
};

}  // namespace hv
