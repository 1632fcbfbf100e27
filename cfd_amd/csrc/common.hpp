// Shared host-side helpers of libcfdgpu: error capture at the C-ABI and small CSR containers.
#pragma once

#include <cfdgpu.h>

#include <cstdint>
#include <memory>
#include <utility>
#include <stdexcept>
#include <string>
#include <vector>

namespace cfdg {

// Exceptions used inside the library; the C-ABI maps them to cfdg_status codes.
struct InvalidArgument : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NoDevice : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct OutOfMemory : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct CapacityError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void set_last_error(const std::string& msg);
const std::string& last_error();

// Runs f and converts any exception into a status plus the per-thread message.
template <class F>
cfdg_status guarded(F&& f) noexcept {
    try {
        f();
        return CFDG_OK;
    } catch (const InvalidArgument& e) {
        set_last_error(e.what());
        return CFDG_ERR_INVALID_ARGUMENT;
    } catch (const NoDevice& e) {
        set_last_error(e.what());
        return CFDG_ERR_NO_DEVICE;
    } catch (const HipError& e) {
        set_last_error(e.what());
        return CFDG_ERR_HIP;
    } catch (const OutOfMemory& e) {
        set_last_error(e.what());
        return CFDG_ERR_OUT_OF_MEMORY;
    } catch (const CapacityError& e) {
        set_last_error(e.what());
        return CFDG_ERR_CAPACITY;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return CFDG_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return CFDG_ERR_INTERNAL;
    } catch (...) {
        set_last_error("unknown error");
        return CFDG_ERR_INTERNAL;
    }
}

inline void require(bool ok, const char* what) {
    if (!ok) {
        throw InvalidArgument(what);
    }
}

// Allocator whose resize leaves new elements default-initialised (indeterminate for int and
// double): large CSR arrays are written in full right after they are sized, and value-initialising
// them first would cost a single-threaded pass over the memory (and its page faults).
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
    template <class U, class... Args>
    void construct(U* p, Args&&... args) {
        ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
    }
    template <class U>
    void construct(U* p) noexcept {
        ::new (static_cast<void*>(p)) U;
    }
};
template <class T>
using RawVec = std::vector<T, NoInitAlloc<T>>;

// Owned CSR matrix (row-major, sorted columns).
struct Csr {
    int n_rows = 0;
    int n_cols = 0;
    RawVec<int> rp{0};
    RawVec<int> ci;
    RawVec<double> va;
    [[nodiscard]] int nnz() const { return static_cast<int>(ci.size()); }
};

// Validates a borrowed CSR view and copies it.
Csr copy_csr(const cfdg_csr& v, const char* what);

// Whether this host thread may fan work out over OpenMP threads (staging.cpp). Threads that
// already run side by side (the server's staging threads) turn it off, so that concurrent calls
// do not each start a full team.
bool host_parallel();
void set_host_parallel(bool on);

// memcpy of a large block split over the host threads (staging.cpp).
void parallel_copy(void* dst, const void* src, std::size_t bytes);

// With CFDG_SETUP_TIMES set, prints "setup <what> level <l>: <ms>" to stderr: the host time since
// the previous call on this thread (the setup's phase breakdown).
void setup_phase(const char* what, int level);

// R (I - A diag(wd)): the restriction of the residual that one damped-Jacobi sweep from zero
// leaves (x1 = diag(wd) b), as one matrix (restriction.cpp).
Csr fused_restriction(const cfdg_csr& r, const cfdg_csr& a, const float* wd);

} // namespace cfdg
