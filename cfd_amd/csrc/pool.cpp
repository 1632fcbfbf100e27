// Device memory pool behind DevBuf.
//
// Buffers are carved out of large hipMalloc slabs (best fit over the free ranges of every slab,
// 256-byte granules) and their ranges go back to the slab on release, where they merge with
// free neighbours. Building and dropping systems therefore costs no hipMalloc or hipFree once
// the pool has grown to the working set; hipFree waits for the whole device, which would stall
// the solves that other threads keep queued. Slabs stay reserved until cfdg_release_cached_memory
// (or an allocation that fails) returns the empty ones to the device.
//
// A range is reused as soon as it is released, so its owner must have no device work pending on
// it by then: every API call ends synchronised, System synchronises its stream before its
// buffers go, and the server releases a system only after the engine stream has passed its last
// use.
#include "hip_util.hpp"

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace cfdg {
namespace {

constexpr std::size_t k_granule = 256;
constexpr std::size_t k_slab = std::size_t{512} << 20;
constexpr std::size_t k_big_granule = std::size_t{2} << 20;

std::size_t round_up(std::size_t v, std::size_t g) { return (v + g - 1) / g * g; }

struct Slab {
    int device = 0;
    char* base = nullptr;
    std::size_t size = 0;
    std::size_t used = 0;
    std::map<std::size_t, std::size_t> free; // offset -> length, disjoint and not adjacent
};

struct Live {
    Slab* slab;
    std::size_t offset;
    std::size_t length;
};

struct Pool {
    std::mutex mu;
    std::vector<std::unique_ptr<Slab>> slabs;
    std::unordered_map<const void*, Live> live;
    std::size_t reserved = 0;
};

// Never destroyed: buffers owned by static objects may be released after static destructors.
Pool& pool() {
    static Pool* p = new Pool;
    return *p;
}

// Frees every empty slab (all devices); the caller holds the lock.
std::size_t release_empty(Pool& p) {
    std::size_t freed = 0;
    auto keep = std::partition(p.slabs.begin(), p.slabs.end(),
                               [](const std::unique_ptr<Slab>& s) { return s->used != 0; });
    for (auto it = keep; it != p.slabs.end(); ++it) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        if ((*it)->device != cur) {
            (void)hipSetDevice((*it)->device);
        }
        (void)hipFree((*it)->base);
        if ((*it)->device != cur) {
            (void)hipSetDevice(cur);
        }
        freed += (*it)->size;
        p.reserved -= (*it)->size;
    }
    p.slabs.erase(keep, p.slabs.end());
    return freed;
}

// Best fit over the free ranges of the device's slabs; null when nothing fits.
void* take(Pool& p, int device, std::size_t len) {
    Slab* best = nullptr;
    std::map<std::size_t, std::size_t>::iterator best_it;
    for (const std::unique_ptr<Slab>& s : p.slabs) {
        if (s->device != device) {
            continue;
        }
        for (auto it = s->free.begin(); it != s->free.end(); ++it) {
            if (it->second >= len && (best == nullptr || it->second < best_it->second)) {
                best = s.get();
                best_it = it;
            }
        }
    }
    if (best == nullptr) {
        return nullptr;
    }
    const std::size_t off = best_it->first;
    const std::size_t rest = best_it->second - len;
    best->free.erase(best_it);
    if (rest > 0) {
        best->free.emplace(off + len, rest);
    }
    best->used += len;
    void* ptr = best->base + off;
    p.live.emplace(ptr, Live{best, off, len});
    return ptr;
}

} // namespace

void* pool_alloc(std::size_t bytes) {
    const std::size_t len = round_up(std::max<std::size_t>(bytes, 1), k_granule);
    int device = 0;
    hip_check(hipGetDevice(&device), "hipGetDevice");
    Pool& p = pool();
    std::lock_guard<std::mutex> lk(p.mu);
    if (void* ptr = take(p, device, len)) {
        return ptr;
    }
    const std::size_t size = len > k_slab / 2 ? round_up(len, k_big_granule) : k_slab;
    void* base = nullptr;
    hipError_t e = hipMalloc(&base, size);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        if (release_empty(p) > 0) {
            e = hipMalloc(&base, size);
        }
    }
    hip_check(e, "hipMalloc (device pool)");
    auto slab = std::make_unique<Slab>();
    slab->device = device;
    slab->base = static_cast<char*>(base);
    slab->size = size;
    slab->free.emplace(0, size);
    p.slabs.push_back(std::move(slab));
    p.reserved += size;
    return take(p, device, len);
}

void pool_free(void* ptr) noexcept {
    if (ptr == nullptr) {
        return;
    }
    Pool& p = pool();
    std::lock_guard<std::mutex> lk(p.mu);
    const auto it = p.live.find(ptr);
    if (it == p.live.end()) {
        return;
    }
    Slab& s = *it->second.slab;
    std::size_t off = it->second.offset;
    std::size_t len = it->second.length;
    p.live.erase(it);
    s.used -= len;
    auto next = s.free.lower_bound(off);
    if (next != s.free.end() && next->first == off + len) { // merge with the range after
        len += next->second;
        next = s.free.erase(next);
    }
    if (next != s.free.begin()) { // merge with the range before
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {
            off = prev->first;
            len += prev->second;
            s.free.erase(prev);
        }
    }
    s.free.emplace(off, len);
}

std::size_t pool_release_cached() {
    Pool& p = pool();
    std::lock_guard<std::mutex> lk(p.mu);
    return release_empty(p);
}

std::size_t pool_reserved_bytes() {
    Pool& p = pool();
    std::lock_guard<std::mutex> lk(p.mu);
    return p.reserved;
}

} // namespace cfdg
