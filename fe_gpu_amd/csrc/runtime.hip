// Device runtime entry points of the drop-in C-ABI (device selection, memory, copies,
// error state).  Replaces the CUDA runtime layer of the reference
// (src/gpu/fe_gpu_runtime.cu:24-70 error state, :449-571 entry points).
#include <cstdarg>
#include <cstring>
#include <vector>

#include "../../include/fe_gpu_runtime.h"
#include "common.hpp"

namespace fehip {

namespace {
thread_local char g_msg[512] = "OK";
std::mutex g_ws_mutex;
struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};
// [device][slot]
std::vector<std::vector<Workspace>> g_ws;
double* g_host_scalars = nullptr;
int g_num_cus = 0;
int g_num_cus_dev = -1;
}  // namespace

int set_ok() {
    std::snprintf(g_msg, sizeof(g_msg), "OK");
    return 0;
}

int set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    return 1;
}

int set_hip_error(hipError_t err, const char* context) {
    if (err == hipSuccess) return set_ok();
    std::snprintf(g_msg, sizeof(g_msg), "%s: %s", context ? context : "HIP", hipGetErrorString(err));
    return static_cast<int>(err);
}

const char* last_error() { return g_msg; }

int num_cus() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (g_num_cus_dev != dev || g_num_cus <= 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
        g_num_cus = v;
        g_num_cus_dev = dev;
    }
    return g_num_cus;
}

void* workspace(size_t bytes, int slot) {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    if ((int)g_ws.size() <= dev) g_ws.resize(dev + 1);
    auto& slots = g_ws[dev];
    if ((int)slots.size() <= slot) slots.resize(slot + 1);
    Workspace& w = slots[slot];
    if (bytes < 256) bytes = 256;
    if (w.bytes >= bytes) return w.ptr;
    if (w.ptr) {
        if (hipDeviceSynchronize() != hipSuccess) return nullptr;
        if (hipFree(w.ptr) != hipSuccess) return nullptr;
        w.ptr = nullptr;
        w.bytes = 0;
    }
    size_t grow = bytes + bytes / 4;
    if (hipMalloc(&w.ptr, grow) != hipSuccess) {
        w.ptr = nullptr;
        return nullptr;
    }
    w.bytes = grow;
    return w.ptr;
}

void release_workspaces() {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) cur = 0;
    for (size_t d = 0; d < g_ws.size(); ++d) {
        if (hipSetDevice((int)d) != hipSuccess) continue;
        for (auto& w : g_ws[d])
            if (w.ptr) {
                (void)hipFree(w.ptr);
                w.ptr = nullptr;
                w.bytes = 0;
            }
    }
    g_ws.clear();
    (void)hipSetDevice(cur);
    if (g_host_scalars) {
        (void)hipHostFree(g_host_scalars);
        g_host_scalars = nullptr;
    }
}

double* host_scalars() {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    if (!g_host_scalars) {
        if (hipHostMalloc((void**)&g_host_scalars, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess)
            g_host_scalars = nullptr;
    }
    return g_host_scalars;
}

// Hook for structures cached per caller pointer (see panel code); invoked by runtime_free.
void forget_pointer(const void* ptr);

}  // namespace fehip

using namespace fehip;

extern "C" {

int fe_gpu_runtime_is_available(void) {
    int count = 0;
    hipError_t err = hipGetDeviceCount(&count);
    if (err != hipSuccess) return set_hip_error(err, "hipGetDeviceCount");
    if (count <= 0) return set_error("No HIP devices detected");
    return set_ok();
}

int fe_gpu_runtime_init(int device_id, fe_gpu_device_info* info) {
    int count = 0;
    FEHIP_CHECK(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (count <= 0) return set_error("No HIP devices detected");
    int chosen = (device_id < 0 || device_id >= count) ? 0 : device_id;
    FEHIP_CHECK(hipSetDevice(chosen), "hipSetDevice");
    hipDeviceProp_t prop;
    FEHIP_CHECK(hipGetDeviceProperties(&prop, chosen), "hipGetDeviceProperties");
    if (info) {
        info->device_id = chosen;
        info->total_global_mem = prop.totalGlobalMem;
        info->multiprocessor_count = prop.multiProcessorCount;
        info->major = prop.major;
        info->minor = prop.minor;
        std::strncpy(info->name, prop.name, FE_GPU_NAME_LEN - 1);
        info->name[FE_GPU_NAME_LEN - 1] = '\0';
    }
    return set_ok();
}

int fe_gpu_runtime_shutdown(void) {
    release_workspaces();
    return set_hip_error(hipDeviceReset(), "hipDeviceReset");
}

int fe_gpu_runtime_malloc(void** ptr, size_t bytes) {
    if (!ptr) return set_error("Null pointer passed to fe_gpu_runtime_malloc");
    if (bytes == 0) {
        *ptr = nullptr;
        return set_ok();
    }
    return set_hip_error(hipMalloc(ptr, bytes), "hipMalloc");
}

int fe_gpu_runtime_free(void* ptr) {
    if (!ptr) return set_ok();
    forget_pointer(ptr);
    return set_hip_error(hipFree(ptr), "hipFree");
}

int fe_gpu_runtime_memcpy_htod(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return set_ok();
    return set_hip_error(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy (host->device)");
}

int fe_gpu_runtime_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return set_ok();
    return set_hip_error(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy (device->host)");
}

int fe_gpu_runtime_memcpy_dtod(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return set_ok();
    return set_hip_error(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy (device->device)");
}

int fe_gpu_runtime_memset(void* ptr, int value, size_t bytes) {
    if (!ptr || bytes == 0) return set_ok();
    return set_hip_error(hipMemset(ptr, value, bytes), "hipMemset");
}

int fe_gpu_runtime_get_last_error(char* buffer, size_t length) {
    if (!buffer || length == 0) return 1;
    size_t n = std::strlen(last_error());
    size_t c = n < length - 1 ? n : length - 1;
    std::memcpy(buffer, last_error(), c);
    buffer[c] = '\0';
    return 0;
}

int fe_gpu_runtime_clear_error(void) {
    (void)hipGetLastError();
    return set_ok();
}

}  // extern "C"
