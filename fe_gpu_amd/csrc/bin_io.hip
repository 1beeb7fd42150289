// fehip_bin_*: the reference's `.bin` dataset file straight into device memory.
//
// Format and header rules follow the reference reader, src/io/fe_data_io.f90:
//   header: int64 N, int32 K, [int32 L], int32 n_fe, int32 has_cluster, int32 has_weights (:7-9, :78-202);
//   the variant with L ("new") or without it ("old") and the value precision are told apart by the file
//   size against compute_data_bytes (:342-358); an int32 precision flag follows the header when the size
//   says so, and any bytes past the data are the META block (:177-201);
//   blocks: y, W (N x K column-major), Z (N x L), FE ids (n_fe x N, dim-major), cluster, weights (:48-62).
// The reference reads every block into host arrays, widens FP32 values to FP64 on the host (:243-299) and
// uploads afterwards (src/gpu/fe_gpu_data.f90:41-120).  Here the file is read in chunks into two pinned
// host buffers (by several reader threads) while the previous chunk is copied to the device on its own
// stream, and FP32 values are widened on the device, so the host never holds the data set.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <thread>

#include "../../include/fehip.h"
#include "common.hpp"

namespace fehip {
namespace {

constexpr int64_t kHeaderNew = 8 + 5 * 4;  // fe_data_io.f90:7
constexpr int64_t kHeaderOld = 8 + 4 * 4;  // fe_data_io.f90:8
constexpr size_t kChunk = (size_t)64 << 20;

// compute_data_bytes (fe_data_io.f90:342-358)
int64_t data_bytes(int64_t n, int32_t K, int32_t L, int32_t n_fe, bool cluster, bool weights, int64_t vb) {
    int64_t t = n * (1 + (int64_t)K) * vb;
    if (L > 0) t += n * (int64_t)L * vb;
    if (weights) t += n * vb;
    t += (int64_t)n_fe * n * 4;
    if (cluster) t += n * 4;
    return t;
}

bool read_exact(int fd, void* dst, size_t bytes, int64_t off);

// a chunk read by several threads at once (reads from the page cache are bound by one core's copy rate)
constexpr int kReaders = 4;
bool read_parallel(int fd, void* dst, size_t bytes, int64_t off) {
    if (bytes < ((size_t)4 << 20)) return read_exact(fd, dst, bytes, off);
    const size_t part = (bytes / kReaders + 4095) & ~(size_t)4095;
    bool ok[kReaders];
    std::thread th[kReaders];
    for (int i = 0; i < kReaders; ++i) {
        const size_t b0 = std::min(bytes, part * (size_t)i), b1 = std::min(bytes, part * (size_t)(i + 1));
        ok[i] = true;
        th[i] = std::thread([&, i, b0, b1] { ok[i] = read_exact(fd, static_cast<char*>(dst) + b0, b1 - b0, off + (int64_t)b0); });
    }
    bool all = true;
    for (int i = 0; i < kReaders; ++i) {
        th[i].join();
        all = all && ok[i];
    }
    return all;
}

bool read_exact(int fd, void* dst, size_t bytes, int64_t off) {
    char* p = static_cast<char*>(dst);
    while (bytes > 0) {
        const ssize_t r = pread(fd, p, bytes, (off_t)off);
        if (r <= 0) return false;
        p += r;
        bytes -= (size_t)r;
        off += r;
    }
    return true;
}

struct Fd {
    int fd = -1;
    ~Fd() {
        if (fd >= 0) close(fd);
    }
};

__global__ void widen_f32_kernel(const float* __restrict__ src, int64_t count, double* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (double)src[i];
}

}  // namespace
}  // namespace fehip

using namespace fehip;

extern "C" int fehip_bin_read_header(const char* path, fehip_bin_header* out) {
    if (!path || !out) return set_error("fehip_bin_read_header: null argument");
    std::memset(out, 0, sizeof(*out));
    Fd f;
    f.fd = open(path, O_RDONLY);
    struct stat st;
    if (f.fd < 0 || fstat(f.fd, &st) != 0) return set_error("Dataset file not found: %s", path);
    const int64_t size = (int64_t)st.st_size;
    if (size < kHeaderOld) return set_error("File too small to contain required header: %s", path);
    unsigned char hb[kHeaderNew + 4];
    std::memset(hb, 0, sizeof(hb));
    if (!read_exact(f.fd, hb, (size_t)std::min<int64_t>(size, (int64_t)sizeof(hb)), 0))
        return set_error("reading header of %s", path);
    int64_t n;
    int32_t v[6];
    std::memcpy(&n, hb, 8);
    std::memcpy(v, hb + 8, 4 * 6);  // K, then four header words, then a possible precision flag
    const int32_t K = v[0];
    if (n <= 0) return set_error("N must be positive");
    if (K < 0) return set_error("K must be non-negative");
    // the layout whose size formula matches the file (fe_data_io.f90:103-125); the new layout otherwise
    auto matches = [&](int64_t base, int32_t L, int32_t nfe, bool cl, bool w) {
        const int64_t t64 = base + data_bytes(n, K, L, nfe, cl, w, 8), t32 = base + data_bytes(n, K, L, nfe, cl, w, 4);
        return size == t64 || size == t32 || size == t64 + 4 || size == t32 + 4;
    };
    const bool mnew = matches(kHeaderNew, v[1], v[2], v[3] != 0, v[4] != 0);
    const bool mold = matches(kHeaderOld, 0, v[1], v[2] != 0, v[3] != 0);
    const bool use_new = mnew || !mold;
    out->n_obs = n;
    out->n_regressors = K;
    out->new_format = use_new ? 1 : 0;
    int64_t base;
    if (use_new) {
        out->n_instruments = v[1];
        out->n_fe = v[2];
        out->has_cluster = v[3] != 0;
        out->has_weights = v[4] != 0;
        base = kHeaderNew;
    } else {
        out->n_instruments = 0;
        out->n_fe = v[1];
        out->has_cluster = v[2] != 0;
        out->has_weights = v[3] != 0;
        base = kHeaderOld;
    }
    if (out->n_instruments < 0) return set_error("L must be non-negative");
    if (out->n_fe < 0) return set_error("n_fe must be non-negative");
    const int64_t d64 = data_bytes(n, K, out->n_instruments, out->n_fe, out->has_cluster, out->has_weights, 8);
    const int64_t d32 = data_bytes(n, K, out->n_instruments, out->n_fe, out->has_cluster, out->has_weights, 4);
    // precision from the size alone, else from the flag word after the header (fe_data_io.f90:173-201)
    if (size == base + d64) {
        out->precision = 0;
        out->data_offset = base;
        return set_ok();
    }
    if (size == base + d32) {
        out->precision = 1;
        out->data_offset = base;
        return set_ok();
    }
    int32_t flag;
    std::memcpy(&flag, hb + base, 4);
    if (flag != 0 && flag != 1) return set_error("Unsupported precision flag value: %d", flag);
    const int64_t expected = base + 4 + (flag == 0 ? d64 : d32);
    if (size < expected) return set_error("File size does not match header metadata for declared precision");
    out->precision = flag;
    out->has_precision_flag = 1;
    out->data_offset = base + 4;
    out->metadata_bytes = size - expected;
    return set_ok();
}

extern "C" int fehip_bin_load(const char* path, const fehip_bin_header* h, double* d_y, double* d_W, double* d_Z,
                              int32_t* d_fe_ids, int32_t* d_cluster, double* d_weights) {
    (void)hipGetLastError();
    if (!path || !h) return set_error("fehip_bin_load: null argument");
    Fd f;
    f.fd = open(path, O_RDONLY);
    if (f.fd < 0) return set_error("Dataset file not found: %s", path);
    const int64_t n = h->n_obs;
    const int64_t vb = h->precision == 1 ? 4 : 8;
    struct Block {
        void* dst;
        int64_t bytes;  // in the file
        bool real;      // numeric block (widened on the device when the file holds FP32)
    };
    // blocks in file order; a null destination skips its bytes
    Block blocks[6] = {{d_y, n * vb, true},
                       {d_W, n * (int64_t)h->n_regressors * vb, true},
                       {d_Z, n * (int64_t)std::max(0, h->n_instruments) * vb, true},
                       {d_fe_ids, n * (int64_t)h->n_fe * 4, false},
                       {d_cluster, h->has_cluster ? n * 4 : 0, false},
                       {d_weights, h->has_weights ? n * vb : 0, true}};
    hipStream_t s = nullptr;
    FEHIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "ingest stream");
    struct Res {
        hipStream_t s;
        void* host[2] = {nullptr, nullptr};
        float* stage[2] = {nullptr, nullptr};
        hipEvent_t done[2] = {nullptr, nullptr};
        ~Res() {
            if (s) (void)hipStreamSynchronize(s);
            for (int i = 0; i < 2; ++i) {
                if (host[i]) (void)locked_host_free(host[i]);
                if (stage[i]) (void)locked_free(stage[i]);
                if (done[i]) (void)hipEventDestroy(done[i]);
            }
            if (s) (void)hipStreamDestroy(s);
        }
    } r{s};
    for (int i = 0; i < 2; ++i) {
        FEHIP_CHECK(hipHostMalloc(&r.host[i], kChunk, hipHostMallocDefault), "pinned staging");
        FEHIP_CHECK(hipEventCreateWithFlags(&r.done[i], hipEventDisableTiming), "ingest event");
        if (vb == 4) FEHIP_CHECK(hipMalloc((void**)&r.stage[i], kChunk), "device staging");
    }
    int64_t off = h->data_offset;
    int slot = 0;
    for (const Block& b : blocks) {
        if (!b.dst) {
            off += b.bytes;
            continue;
        }
        for (int64_t done = 0; done < b.bytes;) {
            const int64_t len = std::min<int64_t>((int64_t)kChunk, b.bytes - done);
            // the pinned buffer (and its device stage) is free once its previous copy has finished
            FEHIP_CHECK(hipEventSynchronize(r.done[slot]), "ingest wait");
            if (!read_parallel(f.fd, r.host[slot], (size_t)len, off + done))
                return set_error("reading %s at byte %lld", path, (long long)(off + done));
            if (b.real && vb == 4) {
                FEHIP_CHECK(hipMemcpyAsync(r.stage[slot], r.host[slot], (size_t)len, hipMemcpyHostToDevice, s), "ingest copy");
                const int64_t cnt = len / 4;
                hipLaunchKernelGGL(widen_f32_kernel, dim3(stream_grid((size_t)cnt, 256)), dim3(256), 0, s,
                                   (const float*)r.stage[slot], cnt, static_cast<double*>(b.dst) + done / 4);
                FEHIP_LAUNCH_CHECK("widen");
            } else {
                FEHIP_CHECK(hipMemcpyAsync(static_cast<char*>(b.dst) + done, r.host[slot], (size_t)len,
                                           hipMemcpyHostToDevice, s),
                            "ingest copy");
            }
            FEHIP_CHECK(hipEventRecord(r.done[slot], s), "ingest event");
            slot ^= 1;
            done += len;
        }
        off += b.bytes;
    }
    FEHIP_CHECK(hipStreamSynchronize(s), "ingest sync");
    return set_ok();
}
