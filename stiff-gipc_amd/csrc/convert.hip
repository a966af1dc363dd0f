// Triplet -> upper block-3x3 BCOO (sorted by (row, col), duplicates summed).
//
// Reference: S/linear_system/utils/converter.cu:23-196 (radix sort of
// row<<32|col keys, run heads, exclusive scan, segmented block reduction with
// FP64 atomics).  Here: counting sort by block row (one pass of row
// histogram + scan + scatter), then a per-row rank sort of (col, triplet
// index) inside one wavefront, and a per-unique-block sequential sum in
// ascending triplet order -- deterministic, no floating-point atomics, and
// each 72-byte triplet value is read exactly once.
//
// Across Newton steps the FEM + mass segment of the triplet store keeps its
// indices (the mesh does not change); only the contact triplets in front of
// it do.  The sort of that static segment is kept as a plan, checked for
// exact index equality on every call, and merged with the freshly sorted
// contact segment.  Every key is still summed in ascending triplet order
// (contact triplets first, then static ones), so the result is identical to
// a full sort.
#include "context.hpp"

#include <algorithm>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

namespace sgipc {
namespace {

constexpr int kSortCap = 256;    // one wavefront per row, LDS: rows up to this many triplets
constexpr int kLongCap = 4096;   // one workgroup per row, LDS: longer rows up to this many;
                                 // longer still: a device radix sort of those rows' triplets
constexpr int kLongThreads = 256;
constexpr int kRowsPerBlock = 4;  // one wavefront per row

__global__ void k_count_rows(long long t0, long long t1, const int* __restrict__ rows, const int* __restrict__ cols,
                             int nbr, int* __restrict__ cnt, int* __restrict__ err) {
    long long t = t0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= t1) return;
    int r = rows[t], c = cols[t];
    if ((unsigned)r >= (unsigned)nbr || (unsigned)c >= (unsigned)nbr) {
        atomicOr(err, 1);
        return;
    }
    atomicAdd(cnt + r, 1);
}

__global__ void k_scatter_rows(long long t0, long long t1, const int* __restrict__ rows, int nbr,
                               const int* __restrict__ start, int* __restrict__ cursor,
                               int* __restrict__ bucket) {
    long long t = t0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= t1) return;
    int r = rows[t];
    if ((unsigned)r >= (unsigned)nbr) return;
    int pos = start[r] + atomicAdd(cursor + r, 1);
    bucket[pos] = (int)t;
}

__device__ inline unsigned long long mkkey(int col, int t) {
    return ((unsigned long long)(unsigned)col << 32) | (unsigned)t;
}

// One wavefront per row: sort the row's triplets by (col, t) and count the
// distinct columns.  sorted[] receives triplet ids in sorted order.  All
// waves of the block reach the same barriers (no early exit).
__global__ void __launch_bounds__(64 * kRowsPerBlock) k_row_sort(int nbr, const int* __restrict__ start,
                                                                 const int* __restrict__ bucket,
                                                                 const int* __restrict__ cols,
                                                                 int* __restrict__ sorted,
                                                                 int* __restrict__ ucount,
                                                                 int* __restrict__ long_rows,
                                                                 int* __restrict__ n_long,
                                                                 int* __restrict__ vlong_rows) {
    __shared__ unsigned long long sk[kRowsPerBlock][kSortCap];
    __shared__ unsigned long long so[kRowsPerBlock][kSortCap];
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int r = blockIdx.x * kRowsPerBlock + w;
    int s = 0, len = 0;
    if (r < nbr) { s = start[r]; len = start[r + 1] - s; }
    const bool fast = len > 0 && len <= kSortCap;
    if (fast)
        for (int i = lane; i < len; i += 64) {
            int t = bucket[s + i];
            sk[w][i] = mkkey(cols[t], t);
        }
    __syncthreads();
    if (fast)
        for (int i = lane; i < len; i += 64) {
            unsigned long long k = sk[w][i];
            int rank = 0;
            for (int j = 0; j < len; ++j) rank += (sk[w][j] < k) ? 1 : 0;
            so[w][rank] = k;
        }
    __syncthreads();
    int heads = 0;
    if (fast) {
        for (int i = lane; i < len; i += 64) {
            unsigned long long k = so[w][i];
            sorted[s + i] = (int)(unsigned)(k & 0xffffffffu);
            bool head = (i == 0) || ((so[w][i - 1] >> 32) != (k >> 32));
            heads += head ? 1 : 0;
        }
    } else if (len > 0 && len <= kLongCap) {
        // long row: sorted by one workgroup in k_row_sort_long
        if (lane == 0) long_rows[atomicAdd(n_long, 1)] = r;
    } else if (len > 0) {
        // very long rows (rows of the affine bodies collect thousands):
        // listed for the radix-sort path (sort_very_long_rows)
        if (lane == 0) vlong_rows[atomicAdd(n_long + 1, 1)] = r;
    }
    heads = wave_sum(heads);
    if (r < nbr && lane == 0 && len <= kSortCap) ucount[r] = heads;
}

// triplet t belongs to a very long row
struct VeryLongRow {
    const int* rows;
    const int* flag;
    int nbr;
    __device__ bool operator()(int t) const {
        const int r = rows[t];
        return (unsigned)r < (unsigned)nbr && flag[r] != 0;
    }
};

__global__ void k_mark_rows(int n, const int* __restrict__ list, int* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[list[i]] = 1;
}

__global__ void k_row_lengths(int n, const int* __restrict__ list, const int* __restrict__ start, int* __restrict__ len) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) len[i] = start[list[i] + 1] - start[list[i]];
}

__global__ void k_row_col_keys(int m, const int* __restrict__ sel, const int* __restrict__ rows,
                               const int* __restrict__ cols, unsigned long long* __restrict__ key) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) key[i] = ((unsigned long long)(unsigned)rows[sel[i]] << 32) | (unsigned)cols[sel[i]];
}

// one workgroup per very long row: its (col, id)-sorted triplet ids into
// place and the count of distinct columns
__global__ void __launch_bounds__(kLongThreads) k_place_rows(int n, const int* __restrict__ list,
                                                             const long long* __restrict__ off,
                                                             const int* __restrict__ start,
                                                             const unsigned long long* __restrict__ key,
                                                             const int* __restrict__ val, int* __restrict__ sorted,
                                                             int* __restrict__ ucount) {
    __shared__ int hsum[kLongThreads / 64];
    for (int q = blockIdx.x; q < n; q += gridDim.x) {
        const int r = list[q];
        const int s = start[r], len = start[r + 1] - s;
        const long long o = off[q];
        int heads = 0;
        for (int i = threadIdx.x; i < len; i += kLongThreads) {
            sorted[s + i] = val[o + i];
            heads += (i == 0 || (unsigned)key[o + i - 1] != (unsigned)key[o + i]) ? 1 : 0;
        }
        heads = wave_sum(heads);
        if ((threadIdx.x & 63) == 0) hsum[threadIdx.x >> 6] = heads;
        __syncthreads();
        if (threadIdx.x == 0) {
            int h = 0;
            for (int w = 0; w < kLongThreads / 64; ++w) h += hsum[w];
            ucount[r] = h;
        }
        __syncthreads();
    }
}

// Rows of kSortCap < len <= kLongCap triplets, one workgroup per row
// (grid-stride over the list): rank sort of (col, triplet id) keys in LDS.
__global__ void __launch_bounds__(kLongThreads) k_row_sort_long(const int* __restrict__ long_rows,
                                                                const int* __restrict__ n_long,
                                                                const int* __restrict__ start,
                                                                const int* __restrict__ bucket,
                                                                const int* __restrict__ cols,
                                                                int* __restrict__ sorted, int* __restrict__ ucount) {
    __shared__ unsigned long long sk[kLongCap];
    __shared__ unsigned long long so[kLongCap];
    __shared__ int hsum[kLongThreads / 64];
    const int nl = *n_long;
    for (int q = blockIdx.x; q < nl; q += gridDim.x) {
        const int r = long_rows[q];
        const int s = start[r], len = start[r + 1] - s;
        for (int i = threadIdx.x; i < len; i += kLongThreads) {
            int t = bucket[s + i];
            sk[i] = mkkey(cols[t], t);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < len; i += kLongThreads) {
            unsigned long long k = sk[i];
            int rank = 0;
            for (int j = 0; j < len; ++j) rank += (sk[j] < k) ? 1 : 0;
            so[rank] = k;
        }
        __syncthreads();
        int heads = 0;
        for (int i = threadIdx.x; i < len; i += kLongThreads) {
            unsigned long long k = so[i];
            sorted[s + i] = (int)(unsigned)(k & 0xffffffffu);
            heads += ((i == 0) || ((so[i - 1] >> 32) != (k >> 32))) ? 1 : 0;
        }
        heads = wave_sum(heads);
        if ((threadIdx.x & 63) == 0) hsum[threadIdx.x >> 6] = heads;
        __syncthreads();
        if (threadIdx.x == 0) {
            int h = 0;
            for (int w = 0; w < kLongThreads / 64; ++w) h += hsum[w];
            ucount[r] = h;
        }
        __syncthreads();
    }
}

// One wavefront per row: each head lane sums its run of duplicates in
// ascending triplet order and writes the unique block.
__global__ void __launch_bounds__(64 * kRowsPerBlock) k_row_reduce(int nbr, const int* __restrict__ start,
                                                                   const int* __restrict__ sorted,
                                                                   const int* __restrict__ cols,
                                                                   const double* __restrict__ vals,
                                                                   const int* __restrict__ ustart,
                                                                   int* __restrict__ orow, int* __restrict__ ocol,
                                                                   double* __restrict__ oval) {
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int r = blockIdx.x * kRowsPerBlock + w;
    if (r >= nbr) return;
    int s = start[r], len = start[r + 1] - s;
    int ubase = ustart[r];
    int seen = 0;
    for (int c0 = 0; c0 < len; c0 += 64) {
        int i = c0 + lane;
        int col = -1;
        bool head = false;
        if (i < len) {
            col = cols[sorted[s + i]];
            head = (i == 0) || (cols[sorted[s + i - 1]] != col);
        }
        unsigned long long hm = __ballot(head);
        if (head) {
            int u = ubase + seen + __popcll(hm & ((1ull << lane) - 1ull));
            double acc[9];
            const double* v0 = vals + 9 * (size_t)sorted[s + i];
#pragma unroll
            for (int e = 0; e < 9; ++e) acc[e] = v0[e];
            for (int j = i + 1; j < len; ++j) {
                int tj = sorted[s + j];
                if (cols[tj] != col) break;
                const double* vj = vals + 9 * (size_t)tj;
#pragma unroll
                for (int e = 0; e < 9; ++e) acc[e] += vj[e];
            }
            orow[u] = r;
            ocol[u] = col;
#pragma unroll
            for (int e = 0; e < 9; ++e) oval[9 * (size_t)u + e] = acc[e];
        }
        seen += __popcll(hm);
    }
}

// unique blocks of a sorted segment: column and first position of each run
__global__ void __launch_bounds__(64 * kRowsPerBlock) k_row_runs(int nbr, const int* __restrict__ start,
                                                                 const int* __restrict__ sorted,
                                                                 const int* __restrict__ cols,
                                                                 const int* __restrict__ ustart,
                                                                 int* __restrict__ ucol, int* __restrict__ urun) {
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int r = blockIdx.x * kRowsPerBlock + w;
    if (r >= nbr) return;
    int s = start[r], len = start[r + 1] - s;
    int ubase = ustart[r], seen = 0;
    for (int c0 = 0; c0 < len; c0 += 64) {
        int i = c0 + lane;
        int col = -1;
        bool head = false;
        if (i < len) {
            col = cols[sorted[s + i]];
            head = (i == 0) || (cols[sorted[s + i - 1]] != col);
        }
        unsigned long long hm = __ballot(head);
        if (head) {
            int u = ubase + seen + __popcll(hm & ((1ull << lane) - 1ull));
            ucol[u] = col;
            urun[u] = s + i;
        }
        seen += __popcll(hm);
    }
}

// flag[0] |= 1 where (a, b) differs from (c, d)
__global__ void k_seg_equal(long long n, const int* __restrict__ a, const int* __restrict__ b,
                            const int* __restrict__ c, const int* __restrict__ d, int* __restrict__ flag) {
    bool diff = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        diff |= (a[i] != c[i]) || (b[i] != d[i]);
    if (__any(diff) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

// run [beg, end) of unique block u of a row: the next unique's run start,
// or the row end for the row's last unique block
__device__ inline int run_end(int u, int uend, const int* urun, int row_end) {
    return (u + 1 < uend) ? urun[u + 1] : row_end;
}

// merged unique count per row (static and contact column lists are each
// sorted and duplicate-free)
__global__ void k_merge_count(int nbr, const int* __restrict__ sus, const int* __restrict__ sucol,
                              const int* __restrict__ dus, const int* __restrict__ ducol, int* __restrict__ mcnt) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nbr) return;
    int i = sus[r], ie = sus[r + 1], j = dus[r], je = dus[r + 1], n = 0;
    while (i < ie || j < je) {
        if (j >= je || (i < ie && sucol[i] < ducol[j])) ++i;
        else if (i >= ie || ducol[j] < sucol[i]) ++j;
        else { ++i; ++j; }
        ++n;
    }
    mcnt[r] = n;
}

// merged blocks of a row: row, column and the two runs {contact begin, len,
// static begin, len} (positions in the segments' sorted id lists)
__global__ void k_merge_write(int nbr, const int* __restrict__ sstart, const int* __restrict__ sus,
                              const int* __restrict__ sucol, const int* __restrict__ surun,
                              const int* __restrict__ dstart, const int* __restrict__ dus,
                              const int* __restrict__ ducol, const int* __restrict__ durun,
                              const int* __restrict__ ustart, int* __restrict__ orow, int* __restrict__ ocol,
                              int4* __restrict__ desc) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nbr) return;
    int i = sus[r], ie = sus[r + 1], j = dus[r], je = dus[r + 1], u = ustart[r];
    while (i < ie || j < je) {
        int4 d = make_int4(0, 0, 0, 0);
        int col = -1;
        bool take_s = i < ie && (j >= je || sucol[i] <= ducol[j]);
        bool take_d = j < je && (i >= ie || ducol[j] <= sucol[i]);
        if (take_d) {
            col = ducol[j];
            d.x = durun[j];
            d.y = run_end(j, je, durun, dstart[r + 1]) - d.x;
            ++j;
        }
        if (take_s) {
            col = sucol[i];
            d.z = surun[i];
            d.w = run_end(i, ie, surun, sstart[r + 1]) - d.z;
            ++i;
        }
        orow[u] = r;
        ocol[u] = col;
        desc[u] = d;
        ++u;
    }
}

// value of each merged block: its contact triplets, then its static ones,
// each run in ascending triplet order
__global__ void k_merge_values(int K, const int4* __restrict__ desc, const int* __restrict__ dsorted,
                               const int* __restrict__ ssorted, int sshift, const double* __restrict__ vals,
                               double* __restrict__ oval) {
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= K) return;
    int4 d = desc[u];
    const int n = d.y + d.w;
    double acc[9];
    // groups of four contributions: their ids, then their values, are loaded
    // before any is added, so the loads overlap; the sum keeps the order
    // q = 0, 1, ... (bit-identical to a sequential loop)
    constexpr int kG = 4;
    for (int q0 = 0; q0 < n; q0 += kG) {
        int t[kG];
#pragma unroll
        for (int j = 0; j < kG; ++j) {
            const int q = q0 + j;
            t[j] = q >= n ? -1 : ((q < d.y) ? dsorted[d.x + q] : ssorted[d.z + (q - d.y)] + sshift);
        }
        double v[kG][9];
#pragma unroll
        for (int j = 0; j < kG; ++j)
            if (t[j] >= 0) {
                const double* src = vals + 9 * (size_t)t[j];
#pragma unroll
                for (int e = 0; e < 9; ++e) v[j][e] = src[e];
            }
#pragma unroll
        for (int j = 0; j < kG; ++j) {
            if (t[j] < 0) break;
            if (q0 + j == 0) {
#pragma unroll
                for (int e = 0; e < 9; ++e) acc[e] = v[j][e];
            } else {
#pragma unroll
                for (int e = 0; e < 9; ++e) acc[e] += v[j][e];
            }
        }
    }
#pragma unroll
    for (int e = 0; e < 9; ++e) oval[9 * (size_t)u + e] = acc[e];
}

__global__ void k_copy_back(int K, const int* __restrict__ r, const int* __restrict__ c,
                            const double* __restrict__ v, int* __restrict__ tr, int* __restrict__ tc,
                            double* __restrict__ tv) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < K) { tr[k] = r[k]; tc[k] = c[k]; }
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; e < 9LL * K; e += stride) tv[e] = v[e];
}

// ---- full symmetric block CSR --------------------------------------------
}  // namespace

// Sort the triplets [t0, t1) by (row, col, triplet id): start (nbr+1 row
// starts into bucket/sorted), sorted (triplet ids), ucnt / ustart (unique
// blocks per row; total -> *total_dev), bad indices -> *err_dev.
// Rows longer than kLongCap (listed by k_row_sort): their triplets picked
// from [t0, t1) in id order, radix-sorted (stably) by (row, col), and placed
// row by row -- the same (col, id) order as the other rows.
static void sort_very_long_rows(Ctx& c, long long t0, long long t1, int nbr, const int* start, int* sorted,
                                int* ucnt, const int* vlong_dev, int nvl) {
    hipStream_t s = c.stream;
    std::vector<int> rows_h(nvl);
    SGIPC_HIP(hipMemcpyAsync(rows_h.data(), vlong_dev, sizeof(int) * nvl, hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    std::sort(rows_h.begin(), rows_h.end());
    int* list = c.vl_rows.reserve((size_t)nvl + 1);
    int* lens = c.vl_lens.reserve((size_t)nvl + 1);
    SGIPC_HIP(hipMemcpyAsync(list, rows_h.data(), sizeof(int) * nvl, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_row_lengths, dim3(div_up(nvl, 256)), dim3(256), 0, s, nvl, (const int*)list, start, lens);
    SGIPC_LAUNCH_CHECK();
    std::vector<int> lens_h(nvl);
    SGIPC_HIP(hipMemcpyAsync(lens_h.data(), lens, sizeof(int) * nvl, hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    std::vector<long long> off_h(nvl);
    long long m = 0;
    for (int i = 0; i < nvl; ++i) { off_h[i] = m; m += lens_h[i]; }
    long long* off = c.vl_off.reserve((size_t)nvl + 1);
    SGIPC_HIP(hipMemcpyAsync(off, off_h.data(), sizeof(long long) * nvl, hipMemcpyHostToDevice, s));
    int* flag = c.vl_flag.reserve((size_t)nbr + 1);
    fill<int>(flag, 0, nbr, s);
    hipLaunchKernelGGL(k_mark_rows, dim3(div_up(nvl, 256)), dim3(256), 0, s, nvl, (const int*)list, flag);
    SGIPC_LAUNCH_CHECK();
    int* sel = c.vl_sel.reserve((size_t)m + 1);
    int* val2 = c.vl_val.reserve((size_t)m + 1);
    unsigned long long* key = c.vl_key.reserve((size_t)m + 1);
    unsigned long long* key2 = c.vl_key2.reserve((size_t)m + 1);
    int* nsel = c.vl_lens.get() + nvl;  // one spare int after the lengths
    rocprim::counting_iterator<int> ids((int)t0);
    VeryLongRow pred{c.t_rows, flag, nbr};
    size_t b1 = 0, b2 = 0;
    SGIPC_HIP(rocprim::select(nullptr, b1, ids, sel, nsel, (size_t)(t1 - t0), pred, s));
    SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, b2, key, key2, sel, val2, (size_t)m, 0, 64, s));
    char* tmp = c.vl_tmp.reserve(std::max(b1, b2) + 1);
    b1 = std::max(b1, b2);
    SGIPC_HIP(rocprim::select((void*)tmp, b1, ids, sel, nsel, (size_t)(t1 - t0), pred, s));
    hipLaunchKernelGGL(k_row_col_keys, dim3(div_up(m, 256)), dim3(256), 0, s, (int)m, (const int*)sel, c.t_rows,
                       c.t_cols, key);
    SGIPC_LAUNCH_CHECK();
    int end_bit = 33;
    while (end_bit < 64 && (1ll << (end_bit - 32)) <= (long long)nbr) ++end_bit;
    b2 = std::max(b1, b2);
    SGIPC_HIP(rocprim::radix_sort_pairs((void*)tmp, b2, key, key2, sel, val2, (size_t)m, 0, end_bit, s));
    hipLaunchKernelGGL(k_place_rows, dim3(std::min(nvl, 1024)), dim3(kLongThreads), 0, s, nvl, (const int*)list,
                       (const long long*)off, start, (const unsigned long long*)key2, (const int*)val2, sorted, ucnt);
    SGIPC_LAUNCH_CHECK();
}

static void sort_segment(Ctx& c, long long t0, long long t1, int nbr, int* cnt, int* start, int* bucket,
                         int* sorted, int* ucnt, int* ustart, int* err_dev, int* total_dev) {
    hipStream_t s = c.stream;
    const long long n = t1 - t0;
    fill<int>(cnt, 0, nbr, s);
    if (n > 0) {
        hipLaunchKernelGGL(k_count_rows, dim3(div_up(n, 256)), dim3(256), 0, s, t0, t1, c.t_rows, c.t_cols, nbr, cnt,
                           err_dev);
        SGIPC_LAUNCH_CHECK();
    }
    exclusive_scan(c, cnt, start, nbr, nullptr, s);
    fill<int>(cnt, 0, nbr, s);
    if (n > 0) {
        hipLaunchKernelGGL(k_scatter_rows, dim3(div_up(n, 256)), dim3(256), 0, s, t0, t1, c.t_rows, nbr, start, cnt,
                           bucket);
        SGIPC_LAUNCH_CHECK();
    }
    int* long_rows = c.long_rows.reserve((size_t)nbr + 1);
    int* vlong_rows = c.vlong_rows.reserve((size_t)nbr + 1);
    int* n_long = c.dev_int2.reserve(2);   // [0] long rows, [1] very long rows
    fill<int>(n_long, 0, 2, s);
    hipLaunchKernelGGL(k_row_sort, dim3(div_up(nbr, kRowsPerBlock)), dim3(64 * kRowsPerBlock), 0, s, nbr, start,
                       bucket, c.t_cols, sorted, ucnt, long_rows, n_long, vlong_rows);
    SGIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_row_sort_long, dim3(1024), dim3(kLongThreads), 0, s, (const int*)long_rows,
                       (const int*)n_long, start, bucket, c.t_cols, sorted, ucnt);
    SGIPC_LAUNCH_CHECK();
    int nvl = 0;
    SGIPC_HIP(hipMemcpyAsync(&nvl, n_long + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    if (nvl > 0) sort_very_long_rows(c, t0, t1, nbr, start, sorted, ucnt, vlong_rows, nvl);
    exclusive_scan(c, ucnt, ustart, nbr, total_dev, s);
}

static void copy_back(Ctx& c, int K, const int* orow, const int* ocol, const double* oval) {
    if (K <= 0) return;
    int blocks = div_up(9LL * K, 256);
    if (blocks > 65536) blocks = 65536;
    if (blocks < div_up(K, 256)) blocks = div_up(K, 256);
    hipLaunchKernelGGL(k_copy_back, dim3(blocks), dim3(256), 0, c.stream, K, orow, ocol, oval, c.t_rows, c.t_cols,
                       c.t_vals);
    SGIPC_LAUNCH_CHECK();
}

int bcoo_convert(Ctx& c, long long nt, int nbr) {
    hipStream_t s = c.stream;
    c.have_bcoo = false;
    c.have_bsr = false;
    if (!c.t_vals) throw state_error("convert: no triplet view set");
    if (nt < 0 || nt > c.t_cap) throw capacity_error("convert: triplet count exceeds capacity");
    if (nbr <= 0) throw arg_error("convert: block row count must be positive");
    if (nt > 0x7fffffffLL) throw arg_error("convert: too many triplets");
    c.nb = nbr;
    int* dint = c.dev_int.reserve(8);
    fill<int>(dint, 0, 8, s);
    int* ustart = c.uniq_start.reserve((size_t)nbr + 2);
    const long long s0 = c.static_begin;
    if (s0 < 0 || s0 >= nt) {
        // ---- one sort of all triplets ----
        int* cnt = c.row_count.reserve((size_t)nbr + 1);
        int* start = c.row_start.reserve((size_t)nbr + 2);
        int* bucket = c.bucket.reserve((size_t)nt + 1);
        int* sorted = c.bcoo_row.reserve((size_t)nt + 1);
        int* ucnt = c.uniq_count.reserve((size_t)nbr + 1);
        sort_segment(c, 0, nt, nbr, cnt, start, bucket, sorted, ucnt, ustart, dint + 2, dint);
        int hk[4] = {0, 0, 0, 0};
        SGIPC_HIP(hipMemcpyAsync(hk, dint, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipStreamSynchronize(s));
        if (hk[2]) throw arg_error("convert: triplet row/col index out of range [0, n_block_rows)");
        const int K = hk[0];
        int* orow = c.bucket.get();  // bucket no longer needed after the sort
        int* ocol = c.bcoo_col.reserve((size_t)K + 1);
        double* oval = c.bcoo_val.reserve(9 * (size_t)K + 1);
        hipLaunchKernelGGL(k_row_reduce, dim3(div_up(nbr, kRowsPerBlock)), dim3(64 * kRowsPerBlock), 0, s, nbr,
                           start, sorted, c.t_cols, c.t_vals, ustart, orow, ocol, oval);
        SGIPC_LAUNCH_CHECK();
        copy_back(c, K, orow, ocol, oval);
        c.K = K;
        c.have_bcoo = true;
        return K;
    }

    // ---- static segment [s0, nt) from its plan, contact segment [0, s0) sorted ----
    ConvertPlan& P = c.cplan;
    const long long ns = nt - s0;
    // the plan holds triplet ids of the segment where it was built; a static
    // segment that moved (the contact count changed) reuses it shifted
    const bool candidate = P.valid && P.ns == ns && P.nbr == nbr;
    if (candidate) {
        int g = (int)std::min<long long>(div_up(ns, 256), 4096);
        hipLaunchKernelGGL(k_seg_equal, dim3(g), dim3(256), 0, s, ns, (const int*)(c.t_rows + s0),
                           (const int*)(c.t_cols + s0), (const int*)P.crow.p, (const int*)P.ccol.p, dint + 5);
        SGIPC_LAUNCH_CHECK();
    }
    int* cnt = c.row_count.reserve((size_t)nbr + 1);
    int* dstart = c.row_start.reserve((size_t)nbr + 2);
    int* dbucket = c.bucket.reserve((size_t)s0 + 1);
    int* dsorted = c.bcoo_row.reserve((size_t)s0 + 1);
    int* ducnt = c.uniq_count.reserve((size_t)nbr + 1);
    int* dus = c.dyn_ustart.reserve((size_t)nbr + 2);
    sort_segment(c, 0, s0, nbr, cnt, dstart, dbucket, dsorted, ducnt, dus, dint + 2, dint);
    int* ducol = c.dyn_ucol.reserve((size_t)s0 + 1);
    int* durun = c.dyn_urun.reserve((size_t)s0 + 1);
    hipLaunchKernelGGL(k_row_runs, dim3(div_up(nbr, kRowsPerBlock)), dim3(64 * kRowsPerBlock), 0, s, nbr, dstart,
                       dsorted, c.t_cols, dus, ducol, durun);
    SGIPC_LAUNCH_CHECK();
    int hk[8] = {0};
    SGIPC_HIP(hipMemcpyAsync(hk, dint, sizeof(int) * 8, hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    if (hk[2]) throw arg_error("convert: triplet row/col index out of range [0, n_block_rows)");
    if (!candidate || hk[5]) {
        // (re)build the plan of the static segment
        P.valid = false;
        int* pstart = P.start.reserve((size_t)nbr + 2);
        int* pbucket = P.bucket.reserve((size_t)ns + 1);
        int* psorted = P.sorted.reserve((size_t)ns + 1);
        int* pucnt = P.ucnt.reserve((size_t)nbr + 1);
        int* pus = P.ustart.reserve((size_t)nbr + 2);
        fill<int>(dint + 6, 0, 2, s);
        sort_segment(c, s0, nt, nbr, cnt, pstart, pbucket, psorted, pucnt, pus, dint + 6, dint + 7);
        int* pucol = P.ucol.reserve((size_t)ns + 1);
        int* purun = P.urun.reserve((size_t)ns + 1);
        hipLaunchKernelGGL(k_row_runs, dim3(div_up(nbr, kRowsPerBlock)), dim3(64 * kRowsPerBlock), 0, s, nbr, pstart,
                           psorted, c.t_cols, pus, pucol, purun);
        SGIPC_LAUNCH_CHECK();
        SGIPC_HIP(hipMemcpyAsync(P.crow.reserve((size_t)ns), c.t_rows + s0, sizeof(int) * ns, hipMemcpyDeviceToDevice,
                                 s));
        SGIPC_HIP(hipMemcpyAsync(P.ccol.reserve((size_t)ns), c.t_cols + s0, sizeof(int) * ns, hipMemcpyDeviceToDevice,
                                 s));
        int h2[2] = {0, 0};
        SGIPC_HIP(hipMemcpyAsync(h2, dint + 6, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipStreamSynchronize(s));
        if (h2[0]) throw arg_error("convert: triplet row/col index out of range [0, n_block_rows)");
        P.Ks = h2[1];
        P.s0 = s0;
        P.ns = ns;
        P.nbr = nbr;
        P.valid = true;
    }
    int* mcnt = c.row_count.get();  // counts no longer needed
    hipLaunchKernelGGL(k_merge_count, dim3(div_up(nbr, 256)), dim3(256), 0, s, nbr, (const int*)P.ustart.p,
                       (const int*)P.ucol.p, (const int*)dus, (const int*)ducol, mcnt);
    SGIPC_LAUNCH_CHECK();
    exclusive_scan(c, mcnt, ustart, nbr, dint + 1, s);
    int K = 0;
    SGIPC_HIP(hipMemcpyAsync(&K, dint + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    int* orow = c.merge_row.reserve((size_t)K + 1);
    int* ocol = c.bcoo_col.reserve((size_t)K + 1);
    int4* desc = c.merge_desc.reserve((size_t)K + 1);
    double* oval = c.bcoo_val.reserve(9 * (size_t)K + 1);
    hipLaunchKernelGGL(k_merge_write, dim3(div_up(nbr, 256)), dim3(256), 0, s, nbr, (const int*)P.start.p,
                       (const int*)P.ustart.p, (const int*)P.ucol.p, (const int*)P.urun.p, (const int*)dstart,
                       (const int*)dus, (const int*)ducol, (const int*)durun, (const int*)ustart, orow, ocol, desc);
    SGIPC_LAUNCH_CHECK();
    if (K > 0) {
        hipLaunchKernelGGL(k_merge_values, dim3(div_up(K, 256)), dim3(256), 0, s, K, (const int4*)desc,
                           (const int*)dsorted, (const int*)P.sorted.p, (int)(s0 - P.s0), (const double*)c.t_vals,
                           oval);
        SGIPC_LAUNCH_CHECK();
    }
    copy_back(c, K, orow, ocol, oval);
    c.K = K;
    c.have_bcoo = true;
    return K;
}

}  // namespace sgipc
