// Large host <-> device copies through pinned staging memory. A copy from or to pageable memory
// runs at a fraction of the PCIe rate; here the pageable side is copied by the host threads into
// (or out of) pinned chunks while the DMA engine moves the previous chunk. Staging areas come
// from a process-wide free list (one per concurrent caller, kept until the process ends).
#include "hip_util.hpp"

#include <omp.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

namespace cfdg {

namespace {

// The host threads of the library's OpenMP regions (LLVM libomp) sleep 1 ms after a region
// instead of the runtime's default 200 ms of spinning, unless the process set KMP_BLOCKTIME or
// OMP_WAIT_POLICY itself. With a CPU quota equal to the worker count (a cgroup cpu.max of 16
// CPUs for 16 threads) the spinning workers and the thread that polls the GPU exceed the quota
// after every host-side setup, the process is throttled for the rest of the scheduler period, and
// solves queued meanwhile stall for milliseconds between two kernels. Runs when the library is
// loaded: kmp_set_blocktime covers the loading thread, kmp_set_defaults threads that start
// OpenMP regions later.
__attribute__((constructor)) void short_omp_blocktime() {
    if (std::getenv("KMP_BLOCKTIME") != nullptr || std::getenv("OMP_WAIT_POLICY") != nullptr) {
        return;
    }
    kmp_set_defaults("KMP_BLOCKTIME=1");
    kmp_set_blocktime(1);
}

constexpr std::size_t k_chunk = std::size_t(16) << 20; // bytes per chunk; two chunks per area

struct Area {
    char* buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    int device = -1;
};

thread_local bool t_host_parallel = true;

std::mutex g_mutex;
std::vector<Area*> g_free;

Area* acquire(int device) {
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        for (std::size_t i = 0; i < g_free.size(); ++i) {
            if (g_free[i]->device == device) {
                Area* a = g_free[i];
                g_free.erase(g_free.begin() + static_cast<std::ptrdiff_t>(i));
                return a;
            }
        }
    }
    auto* a = new Area;
    a->device = device;
    for (int k = 0; k < 2; ++k) {
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&a->buf[k]), k_chunk, hipHostMallocDefault),
                  "hipHostMalloc staging");
        hip_check(hipEventCreateWithFlags(&a->done[k], hipEventDisableTiming), "staging event");
    }
    return a;
}

void release(Area* a) {
    std::lock_guard<std::mutex> lock(g_mutex);
    g_free.push_back(a);
}

// Host memcpy split over the host threads for large blocks.
void host_copy(char* dst, const char* src, std::size_t n) {
    constexpr std::size_t piece = std::size_t(1) << 20;
    const long long pieces = static_cast<long long>((n + piece - 1) / piece);
    if (pieces <= 1 || !t_host_parallel) {
        std::memcpy(dst, src, n);
        return;
    }
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < pieces; ++i) {
        const std::size_t o = static_cast<std::size_t>(i) * piece;
        std::memcpy(dst + o, src + o, n - o < piece ? n - o : piece);
    }
}

} // namespace

bool host_parallel() { return t_host_parallel; }

void set_host_parallel(bool on) { t_host_parallel = on; }

void parallel_copy(void* dst, const void* src, std::size_t bytes) {
    host_copy(static_cast<char*>(dst), static_cast<const char*>(src), bytes);
}

void h2d_staged(void* dst, const void* src, std::size_t bytes, hipStream_t s) {
    if (bytes < k_staged_min_bytes) {
        if (bytes > 0) {
            hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), "H2D copy");
            hip_check(hipStreamSynchronize(s), "H2D copy sync");
        }
        return;
    }
    int device = 0;
    hip_check(hipGetDevice(&device), "hipGetDevice");
    Area* a = acquire(device);
    try {
        const char* in = static_cast<const char*>(src);
        char* out = static_cast<char*>(dst);
        bool used[2] = {false, false};
        int k = 0;
        for (std::size_t o = 0; o < bytes; o += k_chunk, k ^= 1) {
            const std::size_t n = bytes - o < k_chunk ? bytes - o : k_chunk;
            if (used[k]) { // the chunk's previous DMA must be done before it is refilled
                hip_check(hipEventSynchronize(a->done[k]), "staging wait");
            }
            host_copy(a->buf[k], in + o, n);
            hip_check(hipMemcpyAsync(out + o, a->buf[k], n, hipMemcpyHostToDevice, s), "staged H2D");
            hip_check(hipEventRecord(a->done[k], s), "staging record");
            used[k] = true;
        }
        hip_check(hipStreamSynchronize(s), "staged H2D sync");
    } catch (...) {
        (void)hipStreamSynchronize(s);
        release(a);
        throw;
    }
    release(a);
}

void d2h_staged(void* dst, const void* src, std::size_t bytes, hipStream_t s) {
    if (bytes < k_staged_min_bytes) {
        if (bytes > 0) {
            hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s), "D2H copy");
            hip_check(hipStreamSynchronize(s), "D2H copy sync");
        }
        return;
    }
    int device = 0;
    hip_check(hipGetDevice(&device), "hipGetDevice");
    Area* a = acquire(device);
    try {
        const char* in = static_cast<const char*>(src);
        char* out = static_cast<char*>(dst);
        // chunk i goes to buffer i % 2; it is copied out on the host while chunk i + 1 moves
        std::size_t prev_o = 0;
        std::size_t prev_n = 0;
        int k = 0;
        for (std::size_t o = 0; o < bytes; o += k_chunk, k ^= 1) {
            const std::size_t n = bytes - o < k_chunk ? bytes - o : k_chunk;
            hip_check(hipMemcpyAsync(a->buf[k], in + o, n, hipMemcpyDeviceToHost, s), "staged D2H");
            hip_check(hipEventRecord(a->done[k], s), "staging record");
            if (prev_n > 0) {
                hip_check(hipEventSynchronize(a->done[k ^ 1]), "staging wait");
                host_copy(out + prev_o, a->buf[k ^ 1], prev_n);
            }
            prev_o = o;
            prev_n = n;
        }
        if (prev_n > 0) {
            hip_check(hipEventSynchronize(a->done[k ^ 1]), "staging wait");
            host_copy(out + prev_o, a->buf[k ^ 1], prev_n);
        }
    } catch (...) {
        (void)hipStreamSynchronize(s);
        release(a);
        throw;
    }
    release(a);
}

} // namespace cfdg
