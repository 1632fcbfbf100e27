// Library information and error reporting of the C-ABI.
#include "hip_util.hpp"

#include <cstring>
#include <string>

namespace cfdg {

namespace {
thread_local std::string g_last_error;

void copy_out(const std::string& s, char* buf, size_t len) {
    require(buf != nullptr && len > 0, "null or empty output buffer");
    std::strncpy(buf, s.c_str(), len - 1);
    buf[len - 1] = '\0';
}
} // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }
const std::string& last_error() { return g_last_error; }

} // namespace cfdg

using namespace cfdg;

extern "C" int cfdg_version(void) { return CFDG_VERSION_MAJOR * 1000 + CFDG_VERSION_MINOR; }

extern "C" cfdg_status cfdg_last_error(char* buf, size_t len) {
    if (buf == nullptr || len == 0) {
        return CFDG_ERR_INVALID_ARGUMENT;
    }
    std::strncpy(buf, g_last_error.c_str(), len - 1);
    buf[len - 1] = '\0';
    return CFDG_OK;
}

extern "C" cfdg_status cfdg_backend_name(char* buf, size_t len) {
    return guarded([&] { copy_out("hip", buf, len); });
}

extern "C" cfdg_status cfdg_device_available(int* available) {
    return guarded([&] {
        require(available != nullptr, "null argument");
        int count = 0;
        const hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            count = 0;
        }
        *available = count > 0 ? 1 : 0;
    });
}

extern "C" cfdg_status cfdg_release_cached_memory(size_t* freed_bytes, size_t* reserved_bytes) {
    return guarded([&] {
        const std::size_t f = pool_release_cached();
        if (freed_bytes != nullptr) {
            *freed_bytes = f;
        }
        if (reserved_bytes != nullptr) {
            *reserved_bytes = pool_reserved_bytes();
        }
    });
}

extern "C" cfdg_status cfdg_device_description(char* buf, size_t len) {
    return guarded([&] {
        int count = 0;
        const hipError_t e = hipGetDeviceCount(&count);
        std::string s = "hip";
        if (e != hipSuccess || count == 0) {
            (void)hipGetLastError();
            s += " (no HIP device)";
        } else {
            int dev = 0;
            hip_check(hipGetDevice(&dev), "hipGetDevice");
            hipDeviceProp_t prop{};
            hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
            s += std::string(" (device: ") + prop.name + ", " + prop.gcnArchName + ")";
        }
        copy_out(s, buf, len);
    });
}
