// Panel construction: validates and sizes the FE dims (G_d = max id, ids >= 1 as
// src/core/fe_grouping.f90:11-37 requires), orders rows by the primary dim with a stable
// radix sort, and builds the per-segment run structures and cached cluster partitions.
#include <cstring>
#include <cstdlib>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "panel.hpp"

namespace fehip {
namespace {

constexpr int kBlock = 256;

__global__ void minmax_kernel(const int* __restrict__ ids, int64_t n, int* out) {
    int lo = 0x7fffffff, hi = -0x7fffffff;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        int v = ids[i];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        int a = __shfl_xor(lo, off, kWave), b = __shfl_xor(hi, off, kWave);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        atomicMin(&out[0], lo);
        atomicMax(&out[1], hi);
    }
}

__global__ void unsorted_kernel(const int* __restrict__ ids, int64_t n, int* flag) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1; i < n; i += stride)
        bad |= ids[i] < ids[i - 1];
    if (__any(bad) && (threadIdx.x & (kWave - 1)) == 0) atomicOr(flag, 1);
}

__global__ void iota_keys_kernel(const int* __restrict__ ids, int64_t n, unsigned* keys, int* vals) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        keys[i] = (unsigned)(ids[i] - 1);
        vals[i] = (int)i;
    }
}

__global__ void gather_ids_kernel(const int* __restrict__ src, const int* __restrict__ perm, int64_t n, int* dst) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[perm ? perm[i] : i] - 1;
}

__global__ void histogram_kernel(const int* __restrict__ ids, int64_t n, int* counts) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) atomicAdd(&counts[ids[i]], 1);
}

// few groups: per-workgroup LDS counts, one global add per group and workgroup
constexpr int kLdsHistMax = 8192;
__global__ __launch_bounds__(kBlock) void histogram_lds_kernel(const int* __restrict__ ids, int64_t n, int G,
                                                               int* counts) {
    __shared__ int h[kLdsHistMax];
    for (int g = threadIdx.x; g < G; g += blockDim.x) h[g] = 0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) atomicAdd(&h[ids[i]], 1);
    __syncthreads();
    for (int g = threadIdx.x; g < G; g += blockDim.x)
        if (h[g]) atomicAdd(&counts[g], h[g]);
}

__global__ void widen_kernel(const int* __restrict__ c, int64_t n, int64_t* out, int* maxv) {
    int m = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        out[i] = c[i];
        m = c[i] > m ? c[i] : m;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        int o = __shfl_xor(m, off, kWave);
        m = o > m ? o : m;
    }
    if (maxv && (threadIdx.x & (kWave - 1)) == 0) atomicMax(maxv, m);
}

__global__ void set_last_kernel(int64_t* a, int64_t idx, int64_t v) { a[idx] = v; }

// runs of equal id of one dim inside each primary segment
__global__ void count_runs_kernel(const int64_t* __restrict__ seg_off, int64_t G, const int* __restrict__ ids,
                                  int64_t* nruns) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < G; s += stride) {
        int64_t r0 = seg_off[s], r1 = seg_off[s + 1];
        int64_t c = 0;
        int prev = -1;
        for (int64_t r = r0; r < r1; ++r) {
            int v = ids[r];
            c += (r == r0 || v != prev);
            prev = v;
        }
        nruns[s] = c;
    }
}

__global__ void write_runs_kernel(const int64_t* __restrict__ seg_off, int64_t G, const int* __restrict__ ids,
                                  const int64_t* __restrict__ run_off, unsigned* run_grp, int* run_idx, int* run_gid,
                                  int* run_len) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < G; s += stride) {
        int64_t r0 = seg_off[s], r1 = seg_off[s + 1];
        int64_t k = run_off[s];
        int prev = -1;
        int64_t start = r0;
        for (int64_t r = r0; r < r1; ++r) {
            int v = ids[r];
            if (r == r0 || v != prev) {
                if (r > r0) run_len[k - 1] |= (int)(r - start);
                run_grp[k] = (unsigned)v;
                run_idx[k] = (int)k;
                run_gid[k] = v;
                run_len[k] = r == r0 ? (int)0x80000000u : 0;  // bit 31: the segment's first run
                start = r;
                ++k;
            }
            prev = v;
        }
        if (r1 > r0) run_len[k - 1] |= (int)(r1 - start);
    }
}

// Per row: the group-order position of the run the row ends (-1 if the run goes on), so a kernel that
// writes run partials finds each destination with the row's own ids instead of two dependent loads.
__global__ void row_run_dest_kernel(const int64_t* __restrict__ seg_off, int64_t G, const int* __restrict__ ids,
                                    const int64_t* __restrict__ run_off, const int* __restrict__ run_dest,
                                    int* row_rdest) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < G; s += stride) {
        const int64_t r0 = seg_off[s], r1 = seg_off[s + 1];
        int64_t k = run_off[s];
        for (int64_t r = r0; r < r1; ++r) {
            const bool end = r + 1 == r1 || ids[r + 1] != ids[r];
            row_rdest[r] = end ? run_dest[k] : -1;
            if (end) ++k;
        }
    }
}

// dest[order[k]] = k
__global__ void invert_order_kernel(const int* __restrict__ order, int64_t n, int* dest) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        dest[order[k]] = (int)k;
}

__global__ void run_hist_kernel(const unsigned* __restrict__ grp, int64_t n, int* counts) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) atomicAdd(&counts[grp[i]], 1);
}

struct KeyDims {
    const int* ids[4];
    unsigned long long stride[4];
};

__global__ void subset_keys_kernel(KeyDims kd, int nd, int64_t n, unsigned long long* keys, int* rows) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        unsigned long long k = 0;
        for (int j = 0; j < nd; ++j) k += (unsigned long long)kd.ids[j][i] * kd.stride[j];
        keys[i] = k;
        rows[i] = (int)i;
    }
}

// key = hi * glo + lo per row (hi < 2^31, glo <= 2^31: below 2^62), with the row index
__global__ void pair_keys_kernel(const int* __restrict__ hi, unsigned long long glo, const int* __restrict__ lo,
                                 int64_t n, unsigned long long* keys, int* rows) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        keys[i] = (unsigned long long)hi[i] * glo + (unsigned long long)lo[i];
        rows[i] = (int)i;
    }
}

// dense[rows[k]] = cid[k] - 1: the 0-based rank of each row's key
__global__ void scatter_rank_kernel(const int* __restrict__ rows, const int* __restrict__ cid, int64_t n, int* dense) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) dense[rows[k]] = cid[k] - 1;
}

__global__ void heads_kernel(const unsigned long long* __restrict__ k, int64_t n, int* flags) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        flags[i] = (i == 0 || k[i] != k[i - 1]) ? 1 : 0;
}

__global__ void head_offsets_kernel(const int* __restrict__ flags, const int* __restrict__ cid, int64_t n,
                                    int64_t* off) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        if (flags[i]) off[cid[i] - 1] = i;
}

int bits_for(uint64_t maxv) {
    int b = 0;
    while (b < 64 && (maxv >> b) != 0) ++b;
    return b < 1 ? 1 : b;
}

// exclusive scan of int counts[0..G) into int64 off[0..G], off[G] = total; optional max.
int counts_to_offsets(const int* counts, int64_t G, int64_t total, int64_t* off, int* maxv, hipStream_t s) {
    if (G > 0) {
        hipLaunchKernelGGL(widen_kernel, dim3(stream_grid(G, kBlock)), dim3(kBlock), 0, s, counts, G, off, maxv);
        FEHIP_LAUNCH_CHECK("widen kernel");
        size_t tmp = 0;
        FEHIP_CHECK(rocprim::exclusive_scan(nullptr, tmp, off, off, (int64_t)0, (size_t)G, rocprim::plus<int64_t>(), s),
                    "scan size");
        void* t = workspace(tmp, 6);
        if (!t) return set_error("scan workspace");
        FEHIP_CHECK(rocprim::exclusive_scan(t, tmp, off, off, (int64_t)0, (size_t)G, rocprim::plus<int64_t>(), s),
                    "offset scan");
    }
    hipLaunchKernelGGL(set_last_kernel, dim3(1), dim3(1), 0, s, off, G, total);
    FEHIP_LAUNCH_CHECK("set last");
    return 0;
}

void free_panel(Panel* p) {
    if (!p) return;
    // partitions first: the primary-dim partition aliases seg_off
    for (auto& kv : p->partitions) {
        Partition& q = kv.second;
        dfree(q.order);
        dfree(q.pos);
        if (q.off != p->seg_off) dfree(q.off);
        if (q.keys != p->d[p->P].ids) dfree(q.keys);
    }
    p->partitions.clear();
    for (auto& r : p->reg_cols) dfree(r.copy);
    p->reg_cols.clear();
    dfree(p->perm);
    dfree(p->seg_off);
    dfree(p->tile_seg);
    dfree(p->long_segs);
    for (auto& kv : p->diag_cache) dfree(kv.second);
    p->diag_cache.clear();
    for (auto& kv : p->run_couplings) {
        dfree(kv.second.small_rows);
        dfree(kv.second.B);
        dfree(kv.second.B32);
        dfree(kv.second.C);
    }
    p->run_couplings.clear();
    dfree(p->wtile_seg);
    dfree(p->wtile_rec);
    dfree(p->mtile_seg);
    for (int d = 0; d < p->m; ++d) {
        dfree(p->d[d].ids);
        dfree(p->d[d].counts);
        dfree(p->d[d].seg_run_off);
        dfree(p->d[d].run_order);
        dfree(p->d[d].run_dest);
        dfree(p->d[d].row_rdest);
        dfree(p->d[d].grp_run_off);
        dfree(p->d[d].run_gid);
        dfree(p->d[d].run_len);
        dfree(p->d[d].rtile_rec);
        dfree(p->d[d].rtile_seg);
        dfree(p->d[d].rlong_segs);
        dfree(p->d[d].grp_seg);
        dfree(p->d[d].grp_len);
    }
    delete p;
}

int build_runs(Panel* p, int d, hipStream_t s) {
    DimStruct& D = p->d[d];
    const int64_t GP = p->d[p->P].G;
    int st;
    if ((st = dalloc(&D.seg_run_off, GP + 1, &p->bytes))) return st;
    hipLaunchKernelGGL(count_runs_kernel, dim3(stream_grid(GP, kBlock)), dim3(kBlock), 0, s, p->seg_off, GP, D.ids,
                       D.seg_run_off);
    FEHIP_LAUNCH_CHECK("count runs");
    size_t tmp = 0;
    FEHIP_CHECK(rocprim::exclusive_scan(nullptr, tmp, D.seg_run_off, D.seg_run_off, (int64_t)0, (size_t)GP,
                                        rocprim::plus<int64_t>(), s),
                "scan size");
    // total runs = exclusive_scan[GP-1] + count[GP-1]: read the last count before scanning
    int64_t last_count = 0;
    FEHIP_CHECK(hipMemcpyAsync(&last_count, D.seg_run_off + GP - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s),
                "runs copy");
    void* t = workspace(tmp, 6);
    if (!t) return set_error("scan workspace");
    FEHIP_CHECK(rocprim::exclusive_scan(t, tmp, D.seg_run_off, D.seg_run_off, (int64_t)0, (size_t)GP,
                                        rocprim::plus<int64_t>(), s),
                "run offset scan");
    int64_t last_off = 0;
    FEHIP_CHECK(hipMemcpyAsync(&last_off, D.seg_run_off + GP - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s),
                "runs copy");
    FEHIP_CHECK(hipStreamSynchronize(s), "runs sync");
    const int64_t R = last_off + last_count;
    D.n_runs = R;
    hipLaunchKernelGGL(set_last_kernel, dim3(1), dim3(1), 0, s, D.seg_run_off, GP, R);
    // run groups in segment order, then sorted by group (stable)
    unsigned* grp = nullptr;
    unsigned* grp2 = nullptr;
    int* idx2 = nullptr;
    if ((st = dalloc(&grp, R))) return st;
    if ((st = dalloc(&grp2, R))) return st;
    if ((st = dalloc(&D.run_order, R, &p->bytes))) return st;
    if ((st = dalloc(&idx2, R))) return st;
    if ((st = dalloc(&D.run_gid, R, &p->bytes))) return st;
    if ((st = dalloc(&D.run_len, R, &p->bytes))) return st;
    hipLaunchKernelGGL(write_runs_kernel, dim3(stream_grid(GP, kBlock)), dim3(kBlock), 0, s, p->seg_off, GP, D.ids,
                       D.seg_run_off, grp, D.run_order, D.run_gid, D.run_len);
    FEHIP_LAUNCH_CHECK("write runs");
    int* cnt = nullptr;
    if ((st = dalloc(&cnt, (size_t)D.G))) return st;
    FEHIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(int) * (size_t)D.G, s), "memset");
    hipLaunchKernelGGL(run_hist_kernel, dim3(stream_grid(R, kBlock)), dim3(kBlock), 0, s, grp, R, cnt);
    FEHIP_LAUNCH_CHECK("run hist");
    if ((st = dalloc(&D.grp_run_off, (size_t)D.G + 1, &p->bytes))) return st;
    if ((st = counts_to_offsets(cnt, D.G, R, D.grp_run_off, nullptr, s))) return st;
    rocprim::double_buffer<unsigned> kb(grp, grp2);
    rocprim::double_buffer<int> vb(D.run_order, idx2);
    size_t stmp = 0;
    FEHIP_CHECK(rocprim::radix_sort_pairs(nullptr, stmp, kb, vb, (size_t)R, 0, bits_for((uint64_t)D.G), s),
                "sort size");
    void* t2 = workspace(stmp, 7);
    if (!t2) return set_error("sort workspace");
    FEHIP_CHECK(rocprim::radix_sort_pairs(t2, stmp, kb, vb, (size_t)R, 0, bits_for((uint64_t)D.G), s), "run sort");
    if (vb.current() != D.run_order)
        FEHIP_CHECK(hipMemcpyAsync(D.run_order, vb.current(), sizeof(int) * (size_t)R, hipMemcpyDeviceToDevice, s),
                    "run order copy");
    if ((st = dalloc(&D.run_dest, R, &p->bytes))) return st;
    hipLaunchKernelGGL(invert_order_kernel, dim3(stream_grid((size_t)R, kBlock)), dim3(kBlock), 0, s,
                       (const int*)D.run_order, R, D.run_dest);
    FEHIP_LAUNCH_CHECK("run destinations");
    if ((st = dalloc(&D.row_rdest, (size_t)p->n, &p->bytes))) return st;
    hipLaunchKernelGGL(row_run_dest_kernel, dim3(stream_grid(GP, kBlock)), dim3(kBlock), 0, s, p->seg_off, GP, D.ids,
                       (const int64_t*)D.seg_run_off, (const int*)D.run_dest, D.row_rdest);
    FEHIP_LAUNCH_CHECK("row run destinations");
    FEHIP_CHECK(hipStreamSynchronize(s), "runs sync");
    dfree(grp);
    dfree(grp2);
    dfree(idx2);
    dfree(cnt);
    return 0;
}

// Greedy tiles of whole primary segments with at most kTileRows rows; longer segments are listed
// separately.
int build_tiles(Panel* p) {
    const int64_t GP = p->d[p->P].G;
    std::vector<int64_t> off((size_t)GP + 1);
    FEHIP_CHECK(copy_sync(off.data(), p->seg_off, sizeof(int64_t) * (GP + 1), hipMemcpyDeviceToHost), "seg_off copy");
    std::vector<int64_t> tiles, longs;
    tiles.reserve((size_t)(p->n / kTileRows + 16) * 2);
    int64_t s = 0;
    while (s < GP) {
        if (off[s + 1] - off[s] > kTileRows) {
            longs.push_back(s);
            ++s;
            continue;
        }
        int64_t e = s + 1;
        while (e < GP && off[e + 1] - off[s] <= kTileRows && e - s < kTileRows) ++e;
        tiles.push_back(s);
        tiles.push_back(e);
        s = e;
    }
    p->n_tiles = (int64_t)tiles.size() / 2;
    p->n_long = (int64_t)longs.size();
    // operator partition: runs of consecutive short segments (<= kWaveRows rows, empty ones included)
    // packed greedily into wave tiles; runs of consecutive medium segments into block tiles; every tile
    // holds at least one row
    std::vector<int64_t> wt, mt;
    s = 0;
    while (s < GP) {
        const int64_t len = off[s + 1] - off[s];
        if (len > kTileRows) {
            ++s;
            continue;
        }
        const bool short_seg = len <= kWaveRows;
        const int cap = short_seg ? kWaveRows : kTileRows;
        int64_t e = s + 1;
        while (e < GP) {
            const int64_t le = off[e + 1] - off[e];
            if ((le <= kWaveRows) != short_seg || le > kTileRows) break;
            if (off[e + 1] - off[s] > cap || (!short_seg && e - s >= kTileRows)) break;
            ++e;
        }
        auto& v = short_seg ? wt : mt;
        if (off[e] > off[s]) {  // a tile of empty segments only has nothing to process
            v.push_back(s);
            v.push_back(e);
        }
        s = e;
    }
    p->n_wtiles = (int64_t)wt.size() / 2;
    p->n_mtiles = (int64_t)mt.size() / 2;
    std::vector<int64_t> rec((size_t)p->n_wtiles + 1, 0);
    for (int64_t t = 0; t < p->n_wtiles; ++t) {
        const int64_t r0 = off[wt[2 * t]], nr = off[wt[2 * t + 1]] - r0;
        rec[t] = r0 | (nr << 32);
    }
    int st;
    if ((st = dalloc(&p->wtile_seg, wt.size() + 2, &p->bytes)) || (st = dalloc(&p->mtile_seg, mt.size() + 2, &p->bytes)) ||
        (st = dalloc(&p->wtile_rec, rec.size(), &p->bytes)))
        return st;
    FEHIP_CHECK(copy_sync(p->wtile_rec, rec.data(), rec.size() * sizeof(int64_t), hipMemcpyHostToDevice), "wave records");
    if (!wt.empty())
        FEHIP_CHECK(copy_sync(p->wtile_seg, wt.data(), wt.size() * sizeof(int64_t), hipMemcpyHostToDevice), "wave tiles");
    if (!mt.empty())
        FEHIP_CHECK(copy_sync(p->mtile_seg, mt.data(), mt.size() * sizeof(int64_t), hipMemcpyHostToDevice), "block tiles");
    if ((st = dalloc(&p->tile_seg, tiles.size() + 2, &p->bytes)) || (st = dalloc(&p->long_segs, longs.size() + 1, &p->bytes)))
        return st;
    if (!tiles.empty())
        FEHIP_CHECK(copy_sync(p->tile_seg, tiles.data(), tiles.size() * sizeof(int64_t), hipMemcpyHostToDevice), "tiles");
    if (!longs.empty())
        FEHIP_CHECK(copy_sync(p->long_segs, longs.data(), longs.size() * sizeof(int64_t), hipMemcpyHostToDevice),
                    "long segments");
    return 0;
}

}  // namespace

int panel_partition(Panel* p, unsigned mask, const Partition** out, hipStream_t s) {
    std::lock_guard<std::recursive_mutex> guard(p->mu);
    auto it = p->partitions.find(mask);
    if (it != p->partitions.end()) {
        *out = &it->second;
        return 0;
    }
    Partition part;
    if (mask == (1u << p->P)) {
        part.n_clusters = p->d[p->P].G;
        part.order = nullptr;
        part.off = p->seg_off;
        part.keys = p->d[p->P].ids;
    } else {
        // Rows sorted by the subset's mixed-radix key (first dim least significant), clusters = runs of
        // equal keys.  When the key needs more than 64 bits the dims are combined pairwise instead,
        // most significant first: each stage sorts (dense id of the dims so far, next dim's id), a key
        // below 2^62, and replaces the pair by its dense rank; the last stage's order and clusters are
        // those of the single key.
        std::vector<int> dims;
        bool fits = true;
        unsigned long long span = 1;
        for (int d = 0; d < p->m; ++d) {
            if (!(mask & (1u << d))) continue;
            if ((int)dims.size() == 4) return set_error("cluster subsets of more than 4 dims are not supported");
            dims.push_back(d);
            const unsigned long long g = (unsigned long long)std::max(1, p->d[d].G);
            if (span > ~0ull / g) fits = false;
            else span *= g;
        }
        const char* force = getenv("FEHIP_PARTITION_PAIRWISE");  // test hook: the pairwise path for any span
        if (force && force[0] == '1' && dims.size() > 1) fits = false;
        const int64_t n = p->n;
        unsigned long long *k0, *k1;
        int *r0, *r1, *flags, *cid, *dense = nullptr;
        int st;
        if ((st = dalloc(&k0, n))) return st;
        if ((st = dalloc(&k1, n))) return st;
        if ((st = dalloc(&r0, n))) return st;
        if ((st = dalloc(&r1, n))) return st;
        rocprim::double_buffer<unsigned long long> kb(k0, k1);
        rocprim::double_buffer<int> vb(r0, r1);
        // sort the (key, row) pairs over `bits`, flag the heads and number them (cid, 1-based)
        auto rank_keys = [&](int bits, int* nc_out) -> int {
            kb = rocprim::double_buffer<unsigned long long>(k0, k1);
            vb = rocprim::double_buffer<int>(r0, r1);
            size_t tmp = 0;
            FEHIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp, kb, vb, (size_t)n, 0, bits, s), "sort size");
            void* t = workspace(tmp, 7);
            if (!t) return set_error("sort workspace");
            FEHIP_CHECK(rocprim::radix_sort_pairs(t, tmp, kb, vb, (size_t)n, 0, bits, s), "subset sort");
            flags = reinterpret_cast<int*>(kb.alternate());
            cid = flags + n;  // alternate key buffer holds 2n ints
            hipLaunchKernelGGL(heads_kernel, dim3(stream_grid(n, kBlock)), dim3(kBlock), 0, s, kb.current(), n,
                               flags);
            FEHIP_LAUNCH_CHECK("heads");
            size_t tmp2 = 0;
            FEHIP_CHECK(rocprim::inclusive_scan(nullptr, tmp2, flags, cid, (size_t)n, rocprim::plus<int>(), s),
                        "scan");
            void* t2 = workspace(tmp2, 6);
            if (!t2) return set_error("scan workspace");
            FEHIP_CHECK(rocprim::inclusive_scan(t2, tmp2, flags, cid, (size_t)n, rocprim::plus<int>(), s),
                        "cid scan");
            FEHIP_CHECK(hipMemcpyAsync(nc_out, cid + n - 1, sizeof(int), hipMemcpyDeviceToHost, s), "count");
            FEHIP_CHECK(hipStreamSynchronize(s), "partition sync");
            return 0;
        };
        int nc = 0;
        if (fits) {
            KeyDims kd{};
            unsigned long long stride = 1;
            for (size_t j = 0; j < dims.size(); ++j) {
                kd.ids[j] = p->d[dims[j]].ids;
                kd.stride[j] = stride;
                stride *= (unsigned long long)std::max(1, p->d[dims[j]].G);
            }
            hipLaunchKernelGGL(subset_keys_kernel, dim3(stream_grid(n, kBlock)), dim3(kBlock), 0, s, kd,
                               (int)dims.size(), n, k0, r0);
            FEHIP_LAUNCH_CHECK("subset keys");
            if ((st = rank_keys(bits_for(span - 1), &nc))) return st;
        } else {
            if ((st = dalloc(&dense, n))) return st;
            const int* hi = p->d[dims.back()].ids;  // dense id of the dims combined so far (0-based)
            unsigned long long ghi = (unsigned long long)std::max(1, p->d[dims.back()].G);
            for (int j = (int)dims.size() - 2; j >= 0; --j) {
                const int d = dims[j];
                const unsigned long long glo = (unsigned long long)std::max(1, p->d[d].G);
                hipLaunchKernelGGL(pair_keys_kernel, dim3(stream_grid(n, kBlock)), dim3(kBlock), 0, s, hi, glo,
                                   (const int*)p->d[d].ids, n, k0, r0);
                FEHIP_LAUNCH_CHECK("pair keys");
                if ((st = rank_keys(bits_for(ghi * glo - 1), &nc))) return st;
                if (j > 0) {  // the pair's dense rank per row feeds the next stage
                    hipLaunchKernelGGL(scatter_rank_kernel, dim3(stream_grid(n, kBlock)), dim3(kBlock), 0, s,
                                       (const int*)vb.current(), (const int*)cid, n, dense);
                    FEHIP_LAUNCH_CHECK("pair ranks");
                    hi = dense;
                    ghi = (unsigned long long)std::max(1, nc);
                }
            }
        }
        part.n_clusters = nc;
        if ((st = dalloc(&part.off, (size_t)nc + 1, &p->bytes))) return st;
        hipLaunchKernelGGL(head_offsets_kernel, dim3(stream_grid(n, kBlock)), dim3(kBlock), 0, s, flags, cid, n,
                           part.off);
        hipLaunchKernelGGL(set_last_kernel, dim3(1), dim3(1), 0, s, part.off, (int64_t)nc, n);
        FEHIP_LAUNCH_CHECK("offsets");
        if ((st = dalloc(&part.order, (size_t)n, &p->bytes))) return st;
        FEHIP_CHECK(hipMemcpyAsync(part.order, vb.current(), sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, s),
                    "order copy");
        if ((st = dalloc(&part.keys, (size_t)n, &p->bytes))) return st;
        FEHIP_CHECK(hipMemcpyAsync(part.keys, cid, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, s), "keys copy");
        FEHIP_CHECK(hipStreamSynchronize(s), "partition sync");
        dfree(k0);
        dfree(k1);
        dfree(r0);
        dfree(r1);
        if (dense) dfree(dense);
    }
    auto res = p->partitions.emplace(mask, part);
    *out = &res.first->second;
    return 0;
}

int partition_positions(Panel* p, unsigned mask, const int** pos, hipStream_t s) {
    std::lock_guard<std::recursive_mutex> guard(p->mu);
    const Partition* part = nullptr;
    int st = panel_partition(p, mask, &part, s);
    if (st) return st;
    Partition& q = p->partitions.find(mask)->second;
    if (!q.order) {
        *pos = nullptr;
        return 0;
    }
    if (!q.pos) {
        if ((st = dalloc(&q.pos, (size_t)p->n, &p->bytes))) return st;
        hipLaunchKernelGGL(invert_order_kernel, dim3(stream_grid((size_t)p->n, kBlock)), dim3(kBlock), 0, s,
                           (const int*)q.order, p->n, q.pos);
        FEHIP_LAUNCH_CHECK("partition positions");
    }
    *pos = q.pos;
    return 0;
}

namespace {
__global__ void cols_to_f32_kernel(const double* __restrict__ src, int64_t ld, int ncols, int64_t n,
                                   float* __restrict__ dst) {
    const int64_t total = (int64_t)ncols * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / n, i = e - c * n;
        dst[e] = (float)src[c * ld + i];
    }
}

}  // namespace

const float* panel_col32(Panel* p, const double* x) {
    std::lock_guard<std::recursive_mutex> guard(p->mu);
    for (const auto& r : p->reg_cols) {
        if (x < r.base) continue;
        const int64_t off = x - r.base;
        if (off % r.ld != 0 || off / r.ld >= r.ncols) continue;
        return r.copy + (off / r.ld) * p->n;
    }
    return nullptr;
}

}  // namespace fehip

using namespace fehip;

extern "C" {

int fehip_panel_create(const int* const* d_fe_ids, int n_fe, int64_t n_obs, fehip_panel** out) {
    (void)hipGetLastError();  // start from a clean error state
    if (!out) return set_error("fehip_panel_create: null output");
    *out = nullptr;
    if (!d_fe_ids || n_fe <= 0 || n_fe > FEHIP_MAX_FE) return set_error("fehip_panel_create: 1..%d FE dims", FEHIP_MAX_FE);
    if (n_obs <= 0 || n_obs > 0x7fffffffLL) return set_error("fehip_panel_create: n_obs must be in 1..2^31-1");
    hipStream_t s = 0;
    Panel* p = new Panel();
    p->n = n_obs;
    p->m = n_fe;
    (void)hipGetDevice(&p->device);
    int st;
    int* mm = nullptr;
    if ((st = dalloc(&mm, 2))) {
        free_panel(p);
        return st;
    }
    for (int d = 0; d < n_fe; ++d) {
        int init[2] = {0x7fffffff, -0x7fffffff};
        FEHIP_CHECK(copy_sync(mm, init, sizeof(init), hipMemcpyHostToDevice), "init");
        hipLaunchKernelGGL(minmax_kernel, dim3(stream_grid(n_obs, kBlock)), dim3(kBlock), 0, s, d_fe_ids[d], n_obs, mm);
        int h[2];
        FEHIP_CHECK(copy_sync(h, mm, sizeof(h), hipMemcpyDeviceToHost), "minmax");
        if (h[0] < 1) {
            dfree(mm);
            free_panel(p);
            return set_error("FE indices must be positive and contiguous (dimension %d)", d + 1);
        }
        p->d[d].G = h[1];
        p->d[d].small = h[1] <= kSmallMaxGroups;
    }
    p->P = 0;
    for (int d = 1; d < n_fe; ++d)
        if (p->d[d].G > p->d[p->P].G) p->P = d;
    // is the input already ordered by the primary dim?
    int zero = 0, unsorted = 0;
    FEHIP_CHECK(copy_sync(mm, &zero, sizeof(int), hipMemcpyHostToDevice), "init");
    hipLaunchKernelGGL(unsorted_kernel, dim3(stream_grid(n_obs, kBlock)), dim3(kBlock), 0, s, d_fe_ids[p->P], n_obs, mm);
    FEHIP_CHECK(copy_sync(&unsorted, mm, sizeof(int), hipMemcpyDeviceToHost), "sorted flag");
    dfree(mm);
    p->sorted_input = !unsorted;
    if (unsorted) {
        unsigned *k0, *k1;
        int* v1;
        if ((st = dalloc(&k0, n_obs)) || (st = dalloc(&k1, n_obs)) || (st = dalloc(&p->perm, n_obs, &p->bytes)) ||
            (st = dalloc(&v1, n_obs))) {
            free_panel(p);
            return st;
        }
        hipLaunchKernelGGL(iota_keys_kernel, dim3(stream_grid(n_obs, kBlock)), dim3(kBlock), 0, s, d_fe_ids[p->P],
                           n_obs, k0, p->perm);
        rocprim::double_buffer<unsigned> kb(k0, k1);
        rocprim::double_buffer<int> vb(p->perm, v1);
        size_t tmp = 0;
        int bits = bits_for((uint64_t)p->d[p->P].G);
        FEHIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp, kb, vb, (size_t)n_obs, 0, bits, s), "sort size");
        void* t = workspace(tmp, 7);
        if (!t) return set_error("sort workspace");
        FEHIP_CHECK(rocprim::radix_sort_pairs(t, tmp, kb, vb, (size_t)n_obs, 0, bits, s), "primary sort");
        if (vb.current() != p->perm)
            FEHIP_CHECK(hipMemcpyAsync(p->perm, vb.current(), sizeof(int) * (size_t)n_obs, hipMemcpyDeviceToDevice, s),
                        "perm copy");
        FEHIP_CHECK(hipStreamSynchronize(s), "sort sync");
        dfree(k0);
        dfree(k1);
        dfree(v1);
    }
    for (int d = 0; d < n_fe; ++d) {
        DimStruct& D = p->d[d];
        if ((st = dalloc(&D.ids, n_obs, &p->bytes)) || (st = dalloc(&D.counts, (size_t)D.G, &p->bytes))) {
            free_panel(p);
            return st;
        }
        hipLaunchKernelGGL(gather_ids_kernel, dim3(stream_grid(n_obs, kBlock)), dim3(kBlock), 0, s, d_fe_ids[d],
                           p->perm, n_obs, D.ids);
        FEHIP_CHECK(hipMemsetAsync(D.counts, 0, sizeof(int) * (size_t)D.G, s), "memset");
        if (D.G <= kLdsHistMax)
            hipLaunchKernelGGL(histogram_lds_kernel, dim3(stream_grid(n_obs, kBlock, 2)), dim3(kBlock), 0, s, D.ids,
                               n_obs, D.G, D.counts);
        else
            hipLaunchKernelGGL(histogram_kernel, dim3(stream_grid(n_obs, kBlock)), dim3(kBlock), 0, s, D.ids, n_obs,
                               D.counts);
        FEHIP_LAUNCH_CHECK("histogram");
    }
    const int64_t GP = p->d[p->P].G;
    int* maxseg = nullptr;
    if ((st = dalloc(&p->seg_off, GP + 1, &p->bytes)) || (st = dalloc(&maxseg, 1))) {
        free_panel(p);
        return st;
    }
    FEHIP_CHECK(hipMemsetAsync(maxseg, 0, sizeof(int), s), "memset");
    if ((st = counts_to_offsets(p->d[p->P].counts, GP, n_obs, p->seg_off, maxseg, s))) {
        free_panel(p);
        return st;
    }
    FEHIP_CHECK(copy_sync(&p->max_seg, maxseg, sizeof(int), hipMemcpyDeviceToHost), "max seg");
    dfree(maxseg);
    if ((st = build_tiles(p))) {
        free_panel(p);
        return st;
    }
    for (int d = 0; d < n_fe; ++d) {
        if (d == p->P || p->d[d].small) continue;
        if ((st = build_runs(p, d, s))) {
            free_panel(p);
            return st;
        }
    }
    FEHIP_CHECK(locked_device_sync(), "panel build");
    *out = reinterpret_cast<fehip_panel*>(p);
    return set_ok();
}

int fehip_panel_register_columns(fehip_panel* panel, const double* d_cols, int n_cols, int64_t ld) {
    (void)hipGetLastError();
    if (!panel || !d_cols || n_cols < 1 || ld < 1) return set_error("fehip_panel_register_columns: bad argument");
    Panel* p = reinterpret_cast<Panel*>(panel);
    if (ld < p->n) return set_error("fehip_panel_register_columns: leading dim %lld < %lld rows", (long long)ld,
                                    (long long)p->n);
    std::lock_guard<std::recursive_mutex> guard(p->mu);
    // a block registered again (same base and leading dim) is converted again, in place when it fits
    Panel::RegCols* slot = nullptr;
    for (auto& r : p->reg_cols)
        if (r.base == d_cols && r.ld == ld) slot = &r;
    if (slot && slot->ncols < n_cols) {
        dfree(slot->copy);
        p->bytes -= (int64_t)sizeof(float) * slot->ncols * p->n;
        slot->copy = nullptr;
    }
    if (!slot) {
        p->reg_cols.push_back(Panel::RegCols{d_cols, ld, 0, nullptr});
        slot = &p->reg_cols.back();
    }
    int e;
    if (!slot->copy && (e = dalloc(&slot->copy, (size_t)n_cols * (size_t)p->n, &p->bytes))) return e;
    slot->ncols = n_cols;
    hipLaunchKernelGGL(cols_to_f32_kernel, dim3(stream_grid((size_t)n_cols * (size_t)p->n, 256)), dim3(256), 0, 0,
                       d_cols, ld, n_cols, p->n, slot->copy);
    FEHIP_LAUNCH_CHECK("register columns");
    FEHIP_CHECK(hipStreamSynchronize(0), "register columns");
    return set_ok();
}

int fehip_panel_destroy(fehip_panel* panel) {
    if (!panel) return set_ok();
    FEHIP_CHECK(locked_device_sync(), "panel destroy");
    free_panel(reinterpret_cast<Panel*>(panel));
    return set_ok();
}

int fehip_panel_get_info(const fehip_panel* panel, fehip_panel_info* info) {
    if (!panel || !info) return set_error("fehip_panel_get_info: null argument");
    const Panel* p = reinterpret_cast<const Panel*>(panel);
    std::memset(info, 0, sizeof(*info));
    info->n_obs = p->n;
    info->n_fe = p->m;
    info->primary_dim = p->P;
    info->sorted_input = p->sorted_input;
    info->max_segment = p->max_seg;
    for (int d = 0; d < p->m; ++d) {
        info->group_sizes[d] = p->d[d].G;
        info->n_runs[d] = p->d[d].n_runs;
    }
    info->device_bytes = p->bytes;
    return set_ok();
}

}  // extern "C"
