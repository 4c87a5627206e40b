// deviceArray.h -- owning HIP device buffer (drop-in for the reference's
// headers/deviceArray.cuh:12-65: same members).
#ifndef MOSP_DEVICE_ARRAY_H
#define MOSP_DEVICE_ARRAY_H

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <vector>

template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t count) { allocate(count); }
  ~DeviceArray() { (void)hipFree(data_); }
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  DeviceArray(DeviceArray &&o) noexcept : data_(o.data_), size_(o.size_) {
    o.data_ = nullptr;
    o.size_ = 0;
  }
  DeviceArray &operator=(DeviceArray &&o) noexcept {
    if (this != &o) {
      (void)hipFree(data_);
      data_ = o.data_;
      size_ = o.size_;
      o.data_ = nullptr;
      o.size_ = 0;
    }
    return *this;
  }
  // (Re)allocate count elements (at least one); contents undefined.
  bool allocate(size_t count) {
    (void)hipFree(data_);
    data_ = nullptr;
    size_ = count;
    return hipMalloc(reinterpret_cast<void **>(&data_),
                     (count > 0 ? count : 1) * sizeof(T)) == hipSuccess;
  }
  bool upload(const std::vector<T> &host) {
    return allocate(host.size()) &&
           (host.empty() || hipMemcpy(data_, host.data(), host.size() * sizeof(T),
                                      hipMemcpyHostToDevice) == hipSuccess);
  }
  bool download(std::vector<T> &host) const {
    host.resize(size_);
    return size_ == 0 || hipMemcpy(host.data(), data_, size_ * sizeof(T),
                                   hipMemcpyDeviceToHost) == hipSuccess;
  }
  T *data() { return data_; }
  const T *data() const { return data_; }
  size_t size() const { return size_; }

 private:
  T *data_ = nullptr;
  size_t size_ = 0;
};

#endif  // MOSP_DEVICE_ARRAY_H
