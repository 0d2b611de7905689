// Library-wide C-ABI plumbing: thread-local error text, version, device info.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "sw_common.h"

namespace sw {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void clear_error() { g_err[0] = '\0'; }

int device_cu_count() {
    static std::mutex mu;
    static std::vector<int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
    if (cache[dev] == 0) {
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0) cu = 256;
        cache[dev] = cu;
    }
    return cache[dev];
}

}  // namespace sw

extern "C" {

const char* sw_last_error(void) { return sw::g_err; }

int sw_abi_version(void) { return SW_ABI_VERSION; }

int sw_device_info(int* cu_count, char* arch, int arch_len) {
    int dev = 0;
    SW_TRY_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SW_TRY_HIP(hipGetDeviceProperties(&prop, dev));
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (arch && arch_len > 0) {
        std::strncpy(arch, prop.gcnArchName, (size_t)arch_len - 1);
        arch[arch_len - 1] = '\0';
    }
    return SW_OK;
}

}  // extern "C"
