// common.hpp -- HIP error handling and device buffers shared by the library.
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>

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

// Device arrays carved from one allocation: one hipMalloc and one hipFree for a
// whole set of working arrays (a batch solve uses about fifty).  take<T>(n)
// reserves a 256-byte aligned part; commit() allocates them all; a part's get()
// is valid after commit() and until the arena is destroyed.
class DevArena {
 public:
  template <typename T>
  class Part {
   public:
    T* get() const { return a_->mem_.get() ? reinterpret_cast<T*>(a_->mem_.get() + off_) : nullptr; }
    size_t size() const { return n_; }
    void upload(const T* h, size_t n, hipStream_t s)
    {
      if (n) CUOPT_HIP_CHECK(hipMemcpyAsync(get(), h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
    // Elements [off, off + n) from h (h points at element off of the host copy).
    void upload_at(size_t off, const T* h, size_t n, hipStream_t s)
    {
      if (off + n > n_) throw std::out_of_range("DevArena::Part::upload_at past the part's end");
      if (n) CUOPT_HIP_CHECK(hipMemcpyAsync(get() + off, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* h, size_t n, hipStream_t s) const
    {
      if (n) CUOPT_HIP_CHECK(hipMemcpyAsync(h, get(), n * sizeof(T), hipMemcpyDeviceToHost, s));
    }

   private:
    friend class DevArena;
    Part(const DevArena* a, size_t off, size_t n) : a_(a), off_(off), n_(n) {}
    const DevArena* a_;
    size_t off_, n_;
  };
  template <typename T>
  Part<T> take(size_t n)
  {
    if (mem_.get()) throw std::logic_error("DevArena::take after commit");
    const size_t off = bytes_;
    bytes_ += (n * sizeof(T) + 255) & ~(size_t)255;
    return Part<T>(this, off, n);
  }
  void commit() { mem_.resize(std::max<size_t>(bytes_, 1)); }
  size_t bytes() const { return bytes_; }

 private:
  DevBuf<unsigned char> mem_;
  size_t bytes_ = 0;
};

}  // namespace cuopt_amd
