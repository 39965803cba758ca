// common.hpp -- HIP error handling and device buffers shared by the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace cuopt_amd {

struct hip_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define CUOPT_HIP_CHECK(expr)                                                               \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) {                                                                 \
      throw ::cuopt_amd::hip_error(std::string("HIP error '") + hipGetErrorString(_e) +      \
                                   "' at " + __FILE__ + ":" + std::to_string(__LINE__) +    \
                                   " in " #expr);                                           \
    }                                                                                       \
  } while (0)

// Owning device allocation (hipMalloc / hipFree), move-only.
template <typename T>
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t n) { resize(n); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&)            = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept
  {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void resize(size_t n)
  {
    release();
    if (n) CUOPT_HIP_CHECK(hipMalloc(&p_, n * sizeof(T)));
    n_ = n;
  }
  void release()
  {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

  void upload(const T* h, size_t n, hipStream_t s)
  {
    if (n) CUOPT_HIP_CHECK(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
  void download(T* h, size_t n, hipStream_t s) const
  {
    if (n) CUOPT_HIP_CHECK(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  void zero(hipStream_t s)
  {
    if (n_) CUOPT_HIP_CHECK(hipMemsetAsync(p_, 0, n_ * sizeof(T), s));
  }

 private:
  T* p_     = nullptr;
  size_t n_ = 0;
};

}  // namespace cuopt_amd
