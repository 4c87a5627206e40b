// deviceArray.h -- an owning buffer in device memory.
//
// Member set of the reference's DeviceArray (headers/deviceArray.cuh:12-65),
// which host code written against it uses: construct empty or with a count,
// move (not copy), allocate / upload / download returning bool, data(),
// size().  It is built on this library's own allocator and copies
// (mosp_dev_alloc / mosp_dev_free / mosp_memcpy in mosp_hip.h), so a failure
// leaves its message in mosp_last_error(); ownership is a unique_ptr whose
// deleter returns the memory.
#ifndef MOSP_DEVICE_ARRAY_H
#define MOSP_DEVICE_ARRAY_H

#include <cstddef>
#include <memory>
#include <utility>
#include <vector>

#include "mosp_hip.h"

template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t count) { allocate(count); }
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  DeviceArray(DeviceArray &&other) noexcept
      : mem_(std::move(other.mem_)), count_(std::exchange(other.count_, 0)) {}
  DeviceArray &operator=(DeviceArray &&other) noexcept {
    mem_ = std::move(other.mem_);
    count_ = std::exchange(other.count_, 0);
    return *this;
  }

  // A fresh buffer of `count` elements (room for one when count is 0); the
  // old one is returned first.  Contents are unspecified.
  bool allocate(size_t count) {
    mem_.reset();
    count_ = count;
    mem_.reset(static_cast<T *>(mosp_dev_alloc((count == 0 ? 1 : count) * sizeof(T))));
    return mem_ != nullptr;
  }

  // Sized to `host` and filled from it.
  bool upload(const std::vector<T> &host) {
    if (!allocate(host.size())) {
      return false;
    }
    return host.empty() || mosp_memcpy(mem_.get(), host.data(), bytes(), 1) != 0;
  }

  // The whole buffer into `host` (resized to size()).
  bool download(std::vector<T> &host) const {
    host.resize(count_);
    return count_ == 0 || mosp_memcpy(host.data(), mem_.get(), bytes(), 2) != 0;
  }

  T *data() { return mem_.get(); }
  const T *data() const { return mem_.get(); }
  size_t size() const { return count_; }

 private:
  struct Release {
    void operator()(T *p) const noexcept { (void)mosp_dev_free(p); }
  };
  uint64_t bytes() const { return static_cast<uint64_t>(count_) * sizeof(T); }

  std::unique_ptr<T, Release> mem_;
  size_t count_ = 0;
};

#endif  // MOSP_DEVICE_ARRAY_H
