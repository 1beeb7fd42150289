// Shared plumbing for the gfx950 fixed-effects library: error state, launch shapes,
// a grow-only device workspace and small wavefront reduction helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>

namespace fehip {

// ---- error state: one thread-local message, int status (0 = success) ----
int set_ok();
int set_error(const char* fmt, ...);
int set_hip_error(hipError_t err, const char* context);
const char* last_error();

#define FEHIP_CHECK(expr, ctx)                                   \
    do {                                                         \
        hipError_t _e = (expr);                                  \
        if (_e != hipSuccess) return ::fehip::set_hip_error(_e, ctx); \
    } while (0)

#define FEHIP_LAUNCH_CHECK(ctx)                                  \
    do {                                                         \
        hipError_t _e = hipGetLastError();                       \
        if (_e != hipSuccess) return ::fehip::set_hip_error(_e, ctx); \
    } while (0)

// ---- device facts ----
constexpr int kWave = 64;              // CDNA wavefront width
int num_cus();                         // compute units of the current device (256 on MI355X)

// Grid for a grid-stride streaming kernel: enough workgroups to fill every CU several
// times over, never more than the work needs.
inline unsigned stream_grid(size_t work_items, unsigned block, unsigned waves_per_cu = 8) {
    size_t need = (work_items + block - 1) / block;
    size_t fill = (size_t)num_cus() * waves_per_cu * kWave / block;
    if (fill < 1) fill = 1;
    size_t g = need < fill ? need : fill;
    return (unsigned)(g < 1 ? 1 : g);
}

// Workgroups of a grid-stride kernel that the whole device holds at once (its occupancy per CU times
// the CUs).  A larger grid runs in rounds, the last of them partly empty; `need` caps it for small work.
// The occupancy query is remembered per (kernel, block size, LDS bytes).
template <typename F>
unsigned resident_grid(F kernel, int block, size_t lds, size_t need) {
    static std::mutex mu;
    static std::map<std::tuple<const void*, int, size_t>, int> memo;
    const auto key = std::make_tuple(reinterpret_cast<const void*>(kernel), block, lds);
    int per_cu = 0;
    {
        std::lock_guard<std::mutex> l(mu);
        auto it = memo.find(key);
        if (it != memo.end()) per_cu = it->second;
    }
    if (per_cu == 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds) != hipSuccess || per_cu < 1)
            per_cu = 1;
        std::lock_guard<std::mutex> l(mu);
        memo[key] = per_cu;
    }
    const size_t fill = (size_t)per_cu * (size_t)num_cus();
    const size_t g = need < fill ? need : fill;
    return (unsigned)(g < 1 ? 1 : g);
}

// ---- device-synchronising calls and graph capture ----
// hipFree, hipHostFree and hipDeviceSynchronize wait on every stream of the device.  The library runs
// requests from several host threads at once (the request server), and one thread may be capturing a
// CG step into a graph while another frees memory; the capture and these calls therefore exclude each
// other through one process-wide mutex.  Code holding it must not allocate through the caching
// allocator (whose release path takes it).
std::mutex& capture_mutex();
// Whether the calling thread replays the CG step from a captured HIP graph (default) or launches it
// directly; threads that run concurrently with other device work switch capture off.
void set_thread_capture(bool on);
bool thread_capture();  // the calling thread's setting
hipError_t locked_free(void* p);        // hipFree under capture_mutex
hipError_t locked_host_free(void* p);   // hipHostFree under capture_mutex
hipError_t locked_device_sync();        // hipDeviceSynchronize under capture_mutex

// Blocking copy ordered on stream s (the calling thread's default stream unless given).  Used in
// place of hipMemcpy, which the runtime rejects while any thread of the process captures a graph.
inline hipError_t copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s = 0) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
    return e != hipSuccess ? e : hipStreamSynchronize(s);
}

// ---- grow-only workspace on the current device (library-internal temporaries) ----
// Returns a device buffer of at least `bytes`; the pointer stays valid until the next call
// that grows it.  All users run on the default stream, so stream order protects reuse.
void* workspace(size_t bytes, int slot = 0);
void release_workspaces();  // the calling thread's workspaces

// Pinned host scalar buffer for D2H of small results.
double* host_scalars();

// ---- caching allocator for per-call device temporaries ----
// Freed blocks are kept per device and handed out again to requests of a similar size, so a stream
// of regressions does not map and clear fresh device memory for every call.  Reuse is ordered by the
// default stream every user runs on.  cache_release returns the cached blocks to the driver.
void* cache_alloc(size_t bytes);
void cache_free(void* p);
void cache_release();
size_t cache_idle_bytes();  // bytes held by cached (free) blocks on the current device
// Per-thread pinned host scratch (grow-only) for small device-to-host state copies.
void* pinned_scratch(size_t bytes);

// FP64 add into LDS through an explicit LDS-address-space pointer (ds_add_f64); a flat
// pointer would make the backend emit a shared-aperture test.
__device__ __forceinline__ void lds_add(double* p, double v) {
    auto* lp = (__attribute__((address_space(3))) double*)(p);
    __hip_atomic_fetch_add(lp, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void lds_addf(float* p, float v) {
    auto* lp = (__attribute__((address_space(3))) float*)(p);
    __hip_atomic_fetch_add(lp, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Loads through explicit address spaces.  Where one code path reads an LDS table and another a
// global one, the backend may otherwise merge them into one flat load, which waits on both the LDS and
// the vector-memory counters and so serialises the gathers behind unrelated LDS traffic.
__device__ __forceinline__ double ld_global(const double* p) {
    return *(const __attribute__((address_space(1))) double*)(p);
}
__device__ __forceinline__ float ld_global(const float* p) {
    return *(const __attribute__((address_space(1))) float*)(p);
}
__device__ __forceinline__ double ld_lds(const double* p) {
    return *(const __attribute__((address_space(3))) double*)(p);
}

// ---- wavefront helpers ----
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}
// Segmented sum over contiguous lanes: returns the sum of v over lanes [lane, seg_end] where
// seg_end is the last lane of the calling lane's run (after step k a lane holds up to 2^k lanes).
__device__ __forceinline__ double wave_run_sum(double v, int lane, int seg_end) {
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const double t = __shfl_down(v, off, kWave);
        if (lane + off <= seg_end) v += t;
    }
    return v;
}

// Run structure of a per-lane key within the wavefront: returns the last lane of this lane's run
// and whether this lane starts its run.
__device__ __forceinline__ int wave_run_end(int key, int lane, bool* head) {
    const int kn = __shfl_down(key, 1, kWave);
    const int kp = __shfl_up(key, 1, kWave);
    const bool tail = (lane == kWave - 1) || kn != key;
    *head = (lane == 0) || kp != key;
    const unsigned long long above = __ballot(tail) & (~0ull << lane);
    return __ffsll((long long)above) - 1;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}

}  // namespace fehip
