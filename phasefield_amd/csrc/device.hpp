// Device memory, launch and error helpers of the host runtime.
#pragma once

#include <atomic>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

namespace ppf
{
  /// Error raised by the runtime; the C-ABI turns it into a status code.
  struct Error : std::runtime_error
  {
    Error(int code_, const std::string &what)
      : std::runtime_error(what)
      , code(code_)
    {}
    int code;
  };

  constexpr int status_ok             = 0;
  constexpr int status_invalid        = -1;
  constexpr int status_hip            = -2;
  constexpr int status_not_converged  = -3;
  constexpr int status_nan            = -4;
  constexpr int status_unsupported    = -5;
  constexpr int status_internal       = -6;

#define PPF_HIP_CHECK(expr)                                                              \
  do                                                                                     \
    {                                                                                    \
      const hipError_t ppf_err_ = (expr);                                                \
      if (ppf_err_ != hipSuccess)                                                        \
        throw ::ppf::Error(::ppf::status_hip, std::string(#expr) + ": " +                \
                                                hipGetErrorString(ppf_err_));            \
    }                                                                                    \
  while (0)

  /// Owning device allocation of `n` elements of T.
  /// Device allocations made by DeviceBuffer so far (a captured launch sequence is
  /// valid only while no buffer it uses has been reallocated).
  inline std::atomic<long> &
  device_allocations()
  {
    static std::atomic<long> n {0};
    return n;
  }

  template <typename T>
  class DeviceBuffer
  {
  public:
    DeviceBuffer() = default;

    explicit DeviceBuffer(int64_t n)
    {
      resize(n);
    }

    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;

    DeviceBuffer(DeviceBuffer &&o) noexcept
      : ptr_(std::exchange(o.ptr_, nullptr))
      , n_(std::exchange(o.n_, 0))
    {}

    DeviceBuffer &
    operator=(DeviceBuffer &&o) noexcept
    {
      if (this != &o)
        {
          release();
          ptr_ = std::exchange(o.ptr_, nullptr);
          n_   = std::exchange(o.n_, 0);
        }
      return *this;
    }

    ~DeviceBuffer()
    {
      release();
    }

    void
    resize(int64_t n)
    {
      if (n == n_)
        return;
      release();
      if (n > 0)
        {
          PPF_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&ptr_), sizeof(T) * size_t(n)));
          ++device_allocations();
        }
      n_ = n;
    }

    T *
    get() const
    {
      return ptr_;
    }

    int64_t
    size() const
    {
      return n_;
    }

    void
    swap(DeviceBuffer &o) noexcept
    {
      std::swap(ptr_, o.ptr_);
      std::swap(n_, o.n_);
    }

  private:
    void
    release()
    {
      if (ptr_)
        (void) hipFree(ptr_);
      ptr_ = nullptr;
      n_   = 0;
    }

    T      *ptr_ = nullptr;
    int64_t n_   = 0;
  };

  /// Pinned host memory that kernels can write directly (status words read by the host
  /// after a stream synchronization, without a separate copy).
  template <typename T>
  class PinnedBuffer
  {
  public:
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer &) = delete;
    PinnedBuffer &operator=(const PinnedBuffer &) = delete;
    ~PinnedBuffer()
    {
      if (ptr_)
        (void) hipHostFree(ptr_);
    }

    T *
    get(int64_t n = 1)
    {
      if (!ptr_)
        {
          PPF_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&ptr_), sizeof(T) * size_t(n),
                                      hipHostMallocCoherent));
          n_ = n;
        }
      if (n > n_)
        throw Error(status_internal, "pinned buffer too small");
      return ptr_;
    }

  private:
    T      *ptr_ = nullptr;
    int64_t n_   = 0;
  };

  inline unsigned int
  grid_for(int64_t n, int threads)
  {
    return static_cast<unsigned int>((n + threads - 1) / threads);
  }
} // namespace ppf
