// Device-wide exclusive prefix sum of int32 counts: one workgroup for up to
// 16384 entries (the MAS level numbering scans one flag per bank), else three
// passes (tile sums, one-block scan of the tile sums, tile rescan).  Used by
// the BCOO convert, the full-BSR build and the MAS level numbering.
#include "context.hpp"

namespace sgipc {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048
constexpr int kSmallThreads = 1024;
constexpr int kSmallItems = 16;
constexpr int kSmallScan = kSmallThreads * kSmallItems;  // 16384

template <int NT>
__device__ inline int block_exclusive_scan(int v, int* sh, int& total) {
    // sh: NT/64 + 1 ints
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) sh[wid] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < NT / 64; ++w) {
            int t = sh[w];
            sh[w] = acc;
            acc += t;
        }
        sh[NT / 64] = acc;
    }
    __syncthreads();
    total = sh[NT / 64];
    int r = sh[wid] + incl - v;
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kScanThreads) k_tile_sums(const int* __restrict__ in, int n,
                                                            int* __restrict__ tile_sums) {
    __shared__ int sh[kScanThreads / 64 + 1];
    long long base = (long long)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i)
        if (base + i < n) s += in[base + i];
    int total;
    (void)block_exclusive_scan<kScanThreads>(s, sh, total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// single block; ntiles <= kScanTile
__global__ void __launch_bounds__(kScanThreads) k_scan_tile_sums(int* __restrict__ tile_sums, int ntiles,
                                                                 int* __restrict__ total_out) {
    __shared__ int sh[kScanThreads / 64 + 1];
    int vals[kScanItems];
    int base = threadIdx.x * kScanItems, s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        vals[i] = (base + i < ntiles) ? tile_sums[base + i] : 0;
        s += vals[i];
    }
    int total;
    int run = block_exclusive_scan<kScanThreads>(s, sh, total);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < ntiles) tile_sums[base + i] = run;
        run += vals[i];
    }
    if (threadIdx.x == 0 && total_out) *total_out = total;
}

__global__ void __launch_bounds__(kScanThreads) k_tile_apply(const int* __restrict__ in, int n,
                                                             const int* __restrict__ tile_offsets,
                                                             int* __restrict__ out) {
    __shared__ int sh[kScanThreads / 64 + 1];
    long long base = (long long)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int vals[kScanItems], s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        vals[i] = (base + i < n) ? in[base + i] : 0;
        s += vals[i];
    }
    int total;
    int run = block_exclusive_scan<kScanThreads>(s, sh, total) + tile_offsets[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) out[base + i] = run;
        run += vals[i];
    }
    // out has n+1 entries: out[n] = grand total
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) out[n] = run;
}

// n <= kSmallScan: the whole scan in one workgroup.  The input goes through
// LDS so that global loads and stores are coalesced (item k = i * 1024 + t);
// thread t then scans its 16 consecutive items there.  The whole input is
// read before anything is written, so in and out may alias.
__device__ __forceinline__ int small_slot(int k) { return k + (k >> 5); }  // one skip word per 32: the
                                                                            // 16-word lane stride spreads over the banks

__global__ void __launch_bounds__(kSmallThreads) k_scan_small(const int* in, int n, int* out,
                                                              int* __restrict__ total_out) {
    __shared__ int buf[kSmallScan + kSmallScan / 32];
    __shared__ int sh[kSmallThreads / 64 + 1];
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < kSmallItems; ++i) {
        const int k = i * kSmallThreads + t;
        buf[small_slot(k)] = k < n ? in[k] : 0;
    }
    __syncthreads();
    const int base = t * kSmallItems;
    int vals[kSmallItems], s = 0;
#pragma unroll
    for (int i = 0; i < kSmallItems; ++i) {
        vals[i] = buf[small_slot(base + i)];
        s += vals[i];
    }
    int total;
    int run = block_exclusive_scan<kSmallThreads>(s, sh, total);  // ends with a barrier: every read of buf is done
#pragma unroll
    for (int i = 0; i < kSmallItems; ++i) {
        buf[small_slot(base + i)] = run;
        run += vals[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSmallItems; ++i) {
        const int k = i * kSmallThreads + t;
        if (k < n) out[k] = buf[small_slot(k)];
    }
    if (t == 0) {
        out[n] = total;
        if (total_out) *total_out = total;
    }
}

// one level: tile sums of in, their exclusive scan (one block when they fit
// one tile, else the same three passes over them, a level deeper), rescan
static void scan_level(int* scratch, const int* in, int* out, int n, int* total_dev, hipStream_t s) {
    const int ntiles = div_up(n, kScanTile);
    int* tiles = scratch;
    hipLaunchKernelGGL(k_tile_sums, dim3(ntiles), dim3(kScanThreads), 0, s, in, n, tiles);
    SGIPC_LAUNCH_CHECK();
    if (ntiles <= kScanTile) {
        hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(kScanThreads), 0, s, tiles, ntiles, total_dev);
        SGIPC_LAUNCH_CHECK();
    } else {
        scan_level(scratch + ntiles + 1, tiles, tiles, ntiles, total_dev, s);  // in place (aliasing allowed)
    }
    hipLaunchKernelGGL(k_tile_apply, dim3(ntiles), dim3(kScanThreads), 0, s, in, n, tiles, out);
    SGIPC_LAUNCH_CHECK();
}

// out[0..n] (n+1 entries), total_dev optional.  in and out may alias.
void exclusive_scan(Ctx& c, const int* in, int* out, int n, int* total_dev, hipStream_t s) {
    if (n <= 0) {
        fill<int>(out, 0, 1, s);
        if (total_dev) fill<int>(total_dev, 0, 1, s);
        return;
    }
    if (n <= kSmallScan) {
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kSmallThreads), 0, s, in, n, out, total_dev);
        SGIPC_LAUNCH_CHECK();
        return;
    }
    // scratch: the tile sums of every level (each level's scan writes one
    // extra entry, its total)
    size_t need = 0;
    for (long long m = n; m > 0;) {
        const long long t = (m + kScanTile - 1) / kScanTile;
        need += (size_t)t + 1;
        if (t <= kScanTile) break;
        m = t;
    }
    scan_level(c.scan_tmp.reserve(need + 1), in, out, n, total_dev, s);
}

static hipEvent_t pool_get(Ctx& c) {
    if (!c.prof.pool.empty()) {
        hipEvent_t e = c.prof.pool.back();
        c.prof.pool.pop_back();
        return e;
    }
    hipEvent_t e;
    SGIPC_HIP(hipEventCreate(&e));
    return e;
}

void prof_begin(Ctx& c, int slot) {
    if (!c.prof.on) return;
    hipEvent_t a = pool_get(c), b = pool_get(c);
    SGIPC_HIP(hipEventRecord(a, c.stream));
    c.prof.pend[slot].push_back({a, b});
}

void prof_end(Ctx& c, int slot) {
    if (!c.prof.on || c.prof.pend[slot].empty()) return;
    SGIPC_HIP(hipEventRecord(c.prof.pend[slot].back().second, c.stream));
}

void prof_flush(Ctx& c) {
    bool any = false;
    for (int s = 0; s < PROF_NSLOTS; ++s) any |= !c.prof.pend[s].empty();
    if (!any) return;
    SGIPC_HIP(hipStreamSynchronize(c.stream));
    for (int s = 0; s < PROF_NSLOTS; ++s) {
        for (auto& pr : c.prof.pend[s]) {
            float ms = 0.f;
            SGIPC_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
            c.prof.ms[s] += ms;
            c.prof.calls[s] += 1;
            c.prof.pool.push_back(pr.first);
            c.prof.pool.push_back(pr.second);
        }
        c.prof.pend[s].clear();
    }
}

// keep the first `keep` pending records of a slot, recycle the rest (used to
// drop launches that ran after a device-side stop and did no work)
void prof_drop_tail(Ctx& c, int slot, size_t keep) {
    auto& v = c.prof.pend[slot];
    if (v.size() <= keep) return;
    SGIPC_HIP(hipStreamSynchronize(c.stream));
    for (size_t i = keep; i < v.size(); ++i) {
        c.prof.pool.push_back(v[i].first);
        c.prof.pool.push_back(v[i].second);
    }
    v.resize(keep);
}

void prof_release(Ctx& c) {
    for (int s = 0; s < PROF_NSLOTS; ++s) {
        for (auto& pr : c.prof.pend[s]) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
        c.prof.pend[s].clear();
    }
    for (auto e : c.prof.pool) (void)hipEventDestroy(e);
    c.prof.pool.clear();
}

}  // namespace sgipc
