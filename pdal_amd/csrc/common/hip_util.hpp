// Host-side HIP helpers shared by the CP_AL and FA_CP libraries.
//
// Errors from the HIP runtime raise pdal::HipError; every public C entry point
// catches it and returns NULL (the reference exits the process instead,
// LP/CP_AL/internal/utils.h:32-63).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>

namespace pdal {

struct HipError : std::runtime_error {
    explicit HipError(const std::string &m) : std::runtime_error(m) {}
};

inline void hip_check(hipError_t e, const char *what, const char *file, int line)
{
    if (e != hipSuccess) {
        char buf[512];
        snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file,
                 line);
        throw HipError(buf);
    }
}

#define PDAL_HIP(call) ::pdal::hip_check((call), #call, __FILE__, __LINE__)
#define PDAL_LAUNCHED() ::pdal::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

// Owning device buffer (no copy, movable).
template <class T>
class DBuf {
  public:
    DBuf() = default;
    explicit DBuf(size_t n) { alloc(n); }
    ~DBuf() { release(); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
    DBuf &operator=(DBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p_ = std::exchange(o.p_, nullptr);
            n_ = std::exchange(o.n_, 0);
        }
        return *this;
    }
    void alloc(size_t n)
    {
        release();
        n_ = n;
        PDAL_HIP(hipMalloc(reinterpret_cast<void **>(&p_), (n ? n : 1) * sizeof(T)));
    }
    void release()
    {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T *get() const { return p_; }
    size_t size() const { return n_; }
    void upload(const T *h, size_t n, hipStream_t s)
    {
        if (n) PDAL_HIP(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T *h, size_t n, hipStream_t s) const
    {
        if (n) PDAL_HIP(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }

  private:
    T *p_ = nullptr;
    size_t n_ = 0;
};

// Pinned host buffer for small per-block scalar exchange.
template <class T>
class PinnedBuf {
  public:
    explicit PinnedBuf(size_t n = 1) { PDAL_HIP(hipHostMalloc(reinterpret_cast<void **>(&p_), n * sizeof(T))); }
    ~PinnedBuf()
    {
        if (p_) (void)hipHostFree(p_);
    }
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    T *get() const { return p_; }
    T &operator[](size_t i) { return p_[i]; }

  private:
    T *p_ = nullptr;
};

}  // namespace pdal
