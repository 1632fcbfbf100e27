// HIP runtime helpers: error checks that throw, owning device buffers, pinned host buffers.
#pragma once

#include "common.hpp"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <utility>

namespace cfdg {

inline void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) {
        return;
    }
    const std::string msg = std::string("HIP ") + what + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory) {
        throw OutOfMemory(msg);
    }
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) {
        throw NoDevice(msg);
    }
    throw HipError(msg);
}

// Throws NoDevice unless a HIP device is visible.
void require_device();

// Device memory pool (pool.cpp): ranges of large hipMalloc slabs, reused after release without
// hipFree. pool_release_cached returns the empty slabs to the device (bytes freed).
void* pool_alloc(std::size_t bytes);
void pool_free(void* p) noexcept;
std::size_t pool_release_cached();
std::size_t pool_reserved_bytes();

// Synchronous copies between pageable host memory and the device through pinned staging
// (staging.cpp); worth it from a few MiB on.
void h2d_staged(void* dst, const void* src, std::size_t bytes, hipStream_t s);
void d2h_staged(void* dst, const void* src, std::size_t bytes, hipStream_t s);
constexpr std::size_t k_staged_min_bytes = std::size_t(4) << 20;

// Owning device allocation from the pool; the owner must have no device work pending on it
// when it is released or resized.
template <class T>
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(std::size_t n) { resize(n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }

    void resize(std::size_t n) {
        if (n == n_) {
            return;
        }
        release();
        if (n > 0) {
            p_ = static_cast<T*>(pool_alloc(n * sizeof(T)));
        }
        n_ = n;
    }
    void upload(const T* h, std::size_t n, hipStream_t s) {
        resize(n);
        if (n * sizeof(T) >= k_staged_min_bytes) {
            h2d_staged(p_, h, n * sizeof(T), s);
        } else if (n > 0) {
            hip_check(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s), "H2D copy");
            hip_check(hipStreamSynchronize(s), "H2D copy sync");
        }
    }
    void zero(hipStream_t s) {
        if (n_ > 0) {
            hip_check(hipMemsetAsync(p_, 0, n_ * sizeof(T), s), "memset");
        }
    }
    [[nodiscard]] T* get() const { return p_; }
    [[nodiscard]] std::size_t size() const { return n_; }
    [[nodiscard]] std::size_t bytes() const { return n_ * sizeof(T); }

private:
    void release() noexcept {
        if (p_ != nullptr) {
            pool_free(p_);
            p_ = nullptr;
        }
        n_ = 0;
    }
    T* p_ = nullptr;
    std::size_t n_ = 0;
};

// Pinned host allocation (hipHostMalloc), used for small asynchronous read-backs.
template <class T>
class PinnedBuf {
public:
    PinnedBuf() = default;
    explicit PinnedBuf(std::size_t n) {
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T), hipHostMallocDefault),
                  "hipHostMalloc");
        n_ = n;
    }
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    PinnedBuf(PinnedBuf&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    PinnedBuf& operator=(PinnedBuf&& o) noexcept {
        if (this != &o) {
            if (p_ != nullptr) {
                (void)hipHostFree(p_);
            }
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    ~PinnedBuf() {
        if (p_ != nullptr) {
            (void)hipHostFree(p_);
        }
    }
    [[nodiscard]] T* get() const { return p_; }
    [[nodiscard]] std::size_t size() const { return n_; }
    T& operator[](std::size_t i) const { return p_[i]; }

private:
    T* p_ = nullptr;
    std::size_t n_ = 0;
};

// Owning stream at the default priority.
class Stream {
public:
    Stream() { hip_check(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "stream create"); }
    Stream(const Stream&) = delete;
    Stream& operator=(const Stream&) = delete;
    ~Stream() {
        if (s_ != nullptr) {
            (void)hipStreamDestroy(s_);
        }
    }
    [[nodiscard]] hipStream_t get() const { return s_; }

private:
    hipStream_t s_ = nullptr;
};

// Waits for an event by polling it. A blocking wait (hipEventSynchronize) was measured to
// oversleep now and then by about one scheduler tick (~4 ms), leaving the device idle with only
// the next iteration queued; the solve loops poll instead (one host core busy while they wait).
inline void spin_wait(hipEvent_t e, const char* what) {
    for (;;) {
        const hipError_t q = hipEventQuery(e);
        if (q == hipSuccess) {
            return;
        }
        if (q != hipErrorNotReady) {
            hip_check(q, what);
        }
    }
}

class Event {
public:
    Event() { hip_check(hipEventCreate(&e_), "event create"); }
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;
    ~Event() {
        if (e_ != nullptr) {
            (void)hipEventDestroy(e_);
        }
    }
    [[nodiscard]] hipEvent_t get() const { return e_; }

private:
    hipEvent_t e_ = nullptr;
};

} // namespace cfdg
