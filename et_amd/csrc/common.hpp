// Shared helpers for the CCSD residual library: error propagation, device
// buffers, launch checks.  Everything here is host code or trivially inline
// device code; kernels live in the *.hip files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace etg {

// Thrown inside the library; every extern "C" entry point catches it and
// turns it into a non-zero status plus a message for etg_last_error().
struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

[[noreturn]] inline void fail(const std::string& msg) { throw Error(msg); }

inline void check_hip(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess) {
    fail(std::string("HIP error in ") + what + " (" + file + ":" + std::to_string(line) +
         "): " + hipGetErrorString(e));
  }
}

#define ETG_HIP(call) ::etg::check_hip((call), #call, __FILE__, __LINE__)
#define ETG_LAUNCH(what) ::etg::check_hip(hipGetLastError(), what, __FILE__, __LINE__)

// Owning device allocation (bytes); moves only.
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t bytes) { resize(bytes); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_), host_(o.host_) {
    o.ptr_ = nullptr;
    o.bytes_ = 0;
  }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      ptr_ = o.ptr_;
      bytes_ = o.bytes_;
      host_ = o.host_;
      o.ptr_ = nullptr;
      o.bytes_ = 0;
    }
    return *this;
  }
  // Grows the allocation (contents are not preserved).  Work queued on any
  // stream may still use the old block, so the device is drained first.
  // Grows to at least `bytes` (contents are not kept); true when the memory moved.
  bool resize(size_t bytes) { return place(bytes, false); }
  // The same in pinned host memory that kernels read and write over PCIe with
  // the same address (for arrays larger than the free HBM).
  bool resize_host(size_t bytes) { return place(bytes, true); }
  void release() {
    if (ptr_) (void)(host_ ? hipHostFree(ptr_) : hipFree(ptr_));
    ptr_ = nullptr;
    bytes_ = 0;
  }
  template <typename T = double>
  T* as() const { return static_cast<T*>(ptr_); }
  size_t bytes() const { return bytes_; }
  bool on_host() const { return host_; }

 private:
  bool place(size_t bytes, bool host) {
    if (bytes <= bytes_ && host == host_) return false;
    if (ptr_) ETG_HIP(hipDeviceSynchronize());
    release();
    host_ = host;
    if (bytes == 0) return false;
    if (host)
      ETG_HIP(hipHostMalloc(&ptr_, bytes, hipHostMallocDefault));
    else
      ETG_HIP(hipMalloc(&ptr_, bytes));
    bytes_ = bytes;
    return true;
  }
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
  bool host_ = false;
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace etg
