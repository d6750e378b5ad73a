// Bridge from the qlten host API to the C ABI of libqlten_amd
// (include/qlten_amd.h): status -> exception, one lazily initialised device,
// one library stream, and an owning device buffer.
//
// The reference's memory primitives QLMalloc / QLFree / QLMemcpy
// (include/qlten/framework/mem_ops.h:38-104) print and exit on a CUDA error;
// here every failure raises qlten::BackendError with the library's message.
#ifndef QLTEN_AMD_FRAMEWORK_BACKEND_H
#define QLTEN_AMD_FRAMEWORK_BACKEND_H

#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>

#include "qlten_amd.h"

namespace qlten {

class BackendError : public std::runtime_error {
 public:
  BackendError(qlt_status s, const std::string &msg) : std::runtime_error(msg), status(s) {}
  qlt_status status;
};

inline void QltCheck(qlt_status s) {
  if (s != QLT_OK) throw BackendError(s, std::string("libqlten_amd: ") + qlt_last_error());
}

namespace hp_numeric {

// Device used by the host API (environment QLTEN_DEVICE, default 0).
inline int &DeviceOrdinal() {
  static int dev = [] {
    const char *e = std::getenv("QLTEN_DEVICE");
    return e ? std::atoi(e) : 0;
  }();
  return dev;
}

// Initialises the library on first use and returns the stream every host
// API call runs on.  Calls on it are ordered, and every call that hands data
// back to the host synchronises it, so the API behaves as synchronous.
inline void *LibStream() {
  static void *stream = [] {
    QltCheck(qlt_init(DeviceOrdinal()));
    void *s = nullptr;
    QltCheck(qlt_stream_create(&s));
    return s;
  }();
  return stream;
}

inline void SyncLibStream() { QltCheck(qlt_stream_sync(LibStream())); }

// Arithmetic of Contract: native FP64 MFMA (default, the drop-in), or FP64
// emulated by BF16 slices or by int8 digits (both within 1e-12 of FP64).
enum class ContractPrecision { kNativeFP64, kSplitBF16, kSplitInt8 };

inline ContractPrecision &ContractPrecisionSetting() {
  static ContractPrecision p = ContractPrecision::kNativeFP64;
  return p;
}

inline void SetContractPrecision(ContractPrecision p) { ContractPrecisionSetting() = p; }

inline int32_t ContractModeFlag() {
  switch (ContractPrecisionSetting()) {
    case ContractPrecision::kSplitBF16:
      return QLT_MODE_SPLIT_BF16;
    case ContractPrecision::kSplitInt8:
      return QLT_MODE_SPLIT_I8;
    default:
      return QLT_MODE_NATIVE;
  }
}

}  // namespace hp_numeric

// Owning device allocation (the reference keeps pactual_raw_data_ as a raw
// pointer managed by QLMalloc / QLFree).
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t bytes) { Allocate(bytes); }
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  DeviceArray(DeviceArray &&o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) {
    o.ptr_ = nullptr;
    o.bytes_ = 0;
  }
  DeviceArray &operator=(DeviceArray &&o) noexcept {
    if (this != &o) {
      Release();
      ptr_ = o.ptr_;
      bytes_ = o.bytes_;
      o.ptr_ = nullptr;
      o.bytes_ = 0;
    }
    return *this;
  }
  ~DeviceArray() { Release(); }

  void Allocate(size_t bytes) {
    Release();
    hp_numeric::LibStream();
    if (bytes) QltCheck(qlt_malloc(&ptr_, bytes));
    bytes_ = bytes;
  }
  void Release() {
    if (ptr_) {
      hp_numeric::SyncLibStream();
      qlt_free(ptr_);
    }
    ptr_ = nullptr;
    bytes_ = 0;
  }
  void *get() const { return ptr_; }
  size_t bytes() const { return bytes_; }

 private:
  void *ptr_ = nullptr;
  size_t bytes_ = 0;
};

}  // namespace qlten

#endif
