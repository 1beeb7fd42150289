// Device runtime entry points of the drop-in C-ABI (device selection, memory, copies,
// error state).  Replaces the CUDA runtime layer of the reference
// (src/gpu/fe_gpu_runtime.cu:24-70 error state, :449-571 entry points).
#include <cstdarg>
#include <cstring>
#include <map>
#include <unordered_map>
#include <utility>
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
// [device][slot], per host thread (each thread's work runs on its own default stream)
thread_local std::vector<std::vector<Workspace>> g_ws;
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

std::mutex& capture_mutex() {
    static std::mutex m;
    return m;
}

hipError_t locked_free(void* p) {
    if (!p) return hipSuccess;
    std::lock_guard<std::mutex> lock(capture_mutex());
    return hipFree(p);
}

hipError_t locked_host_free(void* p) {
    if (!p) return hipSuccess;
    std::lock_guard<std::mutex> lock(capture_mutex());
    return hipHostFree(p);
}

hipError_t locked_device_sync() {
    std::lock_guard<std::mutex> lock(capture_mutex());
    return hipDeviceSynchronize();
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
        if (hipStreamSynchronize(0) != hipSuccess) return nullptr;
        if (locked_free(w.ptr) != hipSuccess) return nullptr;
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
                (void)locked_free(w.ptr);
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

namespace {
std::mutex g_cache_mutex;
std::multimap<std::pair<int, size_t>, void*> g_cache_free;          // (device, bytes) -> block
std::unordered_map<void*, std::pair<int, size_t>> g_cache_live;      // block -> (device, bytes)

size_t cache_round(size_t b) {
    if (b < 256) b = 256;
    if (b < (1u << 20)) return (b + 511) & ~(size_t)511;
    const size_t g = (size_t)2 << 20;
    return (b + g - 1) & ~(g - 1);
}

void cache_drop_device(int dev) {  // caller holds g_cache_mutex
    for (auto it = g_cache_free.begin(); it != g_cache_free.end();) {
        if (it->first.first == dev) {
            (void)locked_free(it->second);
            it = g_cache_free.erase(it);
        } else {
            ++it;
        }
    }
}
}  // namespace

void* cache_alloc(size_t bytes) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    const size_t sz = cache_round(bytes);
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    auto it = g_cache_free.lower_bound({dev, sz});
    if (it != g_cache_free.end() && it->first.first == dev && it->first.second <= sz + sz / 4 + ((size_t)4 << 20)) {
        void* p = it->second;
        g_cache_live[p] = it->first;
        g_cache_free.erase(it);
        return p;
    }
    void* p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        cache_drop_device(dev);  // give the cached blocks back and try once more
        if (hipMalloc(&p, sz) != hipSuccess) return nullptr;
    }
    g_cache_live[p] = {dev, sz};
    return p;
}

void cache_free(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    auto it = g_cache_live.find(p);
    if (it == g_cache_live.end()) return;
    g_cache_free.emplace(it->second, p);
    g_cache_live.erase(it);
}

void cache_release() {
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) cur = 0;
    for (auto& kv : g_cache_free) {
        (void)hipSetDevice(kv.first.first);
        (void)locked_free(kv.second);
    }
    g_cache_free.clear();
    (void)hipSetDevice(cur);
}

size_t cache_idle_bytes() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    size_t b = 0;
    for (auto& kv : g_cache_free)
        if (kv.first.first == dev) b += kv.first.second;
    return b;
}

void* pinned_scratch(size_t bytes) {
    struct Pinned {
        void* p = nullptr;
        size_t bytes = 0;
        ~Pinned() {
            if (p) (void)locked_host_free(p);
        }
    };
    thread_local Pinned buf;
    if (buf.bytes >= bytes) return buf.p;
    if (buf.p) (void)locked_host_free(buf.p);
    buf.p = nullptr;
    buf.bytes = 0;
    if (hipHostMalloc(&buf.p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    buf.bytes = bytes;
    return buf.p;
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
        // the marketing name can be empty (it is on MI355X): fall back to the architecture name (gfx950...)
        const char* name = prop.name[0] ? prop.name : prop.gcnArchName;
        std::strncpy(info->name, name, FE_GPU_NAME_LEN - 1);
        info->name[FE_GPU_NAME_LEN - 1] = '\0';
    }
    return set_ok();
}

int fe_gpu_runtime_shutdown(void) {
    release_workspaces();
    cache_release();
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
