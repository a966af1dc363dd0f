// Triplet -> upper block-3x3 BCOO (sorted by (row, col), duplicates summed)
// and the full symmetric block-CSR used by the SpMV.
//
// Reference: S/linear_system/utils/converter.cu:23-196 (radix sort of
// row<<32|col keys, run heads, exclusive scan, segmented block reduction with
// FP64 atomics).  Here: counting sort by block row (one pass of row
// histogram + scan + scatter), then a per-row rank sort of (col, triplet
// index) inside one wavefront, and a per-unique-block sequential sum in
// ascending triplet order -- deterministic, no floating-point atomics, and
// each 72-byte triplet value is read exactly once.
#include "context.hpp"

namespace sgipc {
namespace {

constexpr int kSortCap = 256;  // LDS fast path: rows up to this many triplets
constexpr int kRowsPerBlock = 4;  // one wavefront per row

__global__ void k_count_rows(long long nt, const int* __restrict__ rows, const int* __restrict__ cols,
                             int nbr, int* __restrict__ cnt, int* __restrict__ err) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int r = rows[t], c = cols[t];
    if ((unsigned)r >= (unsigned)nbr || (unsigned)c >= (unsigned)nbr) {
        atomicOr(err, 1);
        return;
    }
    atomicAdd(cnt + r, 1);
}

__global__ void k_scatter_rows(long long nt, const int* __restrict__ rows, int nbr,
                               const int* __restrict__ start, int* __restrict__ cursor,
                               int* __restrict__ bucket) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
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
                                                                 int* __restrict__ ucount) {
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
    } else if (len > 0) {
        // long rows (rare): rank counting straight from global memory
        for (int i = lane; i < len; i += 64) {
            int ti = bucket[s + i];
            unsigned long long k = mkkey(cols[ti], ti);
            int rank = 0;
            bool dup_before = false;  // same col with a smaller triplet id
            for (int j = 0; j < len; ++j) {
                int tj = bucket[s + j];
                unsigned long long kj = mkkey(cols[tj], tj);
                rank += (kj < k) ? 1 : 0;
                dup_before |= ((kj >> 32) == (k >> 32)) && (kj < k);
            }
            sorted[s + rank] = ti;
            heads += dup_before ? 0 : 1;
        }
    }
    heads = wave_sum(heads);
    if (r < nbr && lane == 0) ucount[r] = heads;
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

int bcoo_convert(Ctx& c, long long nt, int nbr) {
    hipStream_t s = c.stream;
    c.have_bcoo = false;
    c.have_bsr = false;
    if (!c.t_vals) throw state_error("convert: no triplet view set");
    if (nt < 0 || nt > c.t_cap) throw capacity_error("convert: triplet count exceeds capacity");
    if (nbr <= 0) throw arg_error("convert: block row count must be positive");
    if (nt > 0x7fffffffLL) throw arg_error("convert: too many triplets");
    c.nb = nbr;
    int* cnt = c.row_count.reserve((size_t)nbr + 1);
    int* start = c.row_start.reserve((size_t)nbr + 2);
    int* bucket = c.bucket.reserve((size_t)nt + 1);
    int* sorted = c.bcoo_row.reserve((size_t)nt + 1);  // reused below as sort output first
    int* ucnt = c.uniq_count.reserve((size_t)nbr + 1);
    int* ustart = c.uniq_start.reserve((size_t)nbr + 2);
    int* dint = c.dev_int.reserve(8);
    fill<int>(cnt, 0, nbr, s);
    fill<int>(dint, 0, 8, s);
    if (nt > 0) {
        hipLaunchKernelGGL(k_count_rows, dim3(div_up(nt, 256)), dim3(256), 0, s, nt, c.t_rows, c.t_cols, nbr, cnt,
                           dint + 2);
        SGIPC_LAUNCH_CHECK();
    }
    exclusive_scan(c, cnt, start, nbr, nullptr, s);
    fill<int>(cnt, 0, nbr, s);
    if (nt > 0) {
        hipLaunchKernelGGL(k_scatter_rows, dim3(div_up(nt, 256)), dim3(256), 0, s, nt, c.t_rows, nbr, start, cnt,
                           bucket);
        SGIPC_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_row_sort, dim3(div_up(nbr, kRowsPerBlock)), dim3(64 * kRowsPerBlock), 0, s, nbr, start,
                       bucket, c.t_cols, sorted, ucnt);
    SGIPC_LAUNCH_CHECK();
    exclusive_scan(c, ucnt, ustart, nbr, dint, s);
    int hk[4] = {0, 0, 0, 0};
    SGIPC_HIP(hipMemcpyAsync(hk, dint, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    if (hk[2]) throw arg_error("convert: triplet row/col index out of range [0, n_block_rows)");
    int K = hk[0];
    // output: scratch BCOO then copy back over the triplet view (reference
    // leaves the BCOO in place at offset 0)
    int* orow = c.bucket.get();  // bucket no longer needed after sort
    if (c.bucket.cap < (size_t)K + 1) throw state_error("convert: internal sizing");
    int* ocol = c.bcoo_col.reserve((size_t)K + 1);
    double* oval = c.bcoo_val.reserve(9 * (size_t)K + 1);
    hipLaunchKernelGGL(k_row_reduce, dim3(div_up(nbr, kRowsPerBlock)), dim3(64 * kRowsPerBlock), 0, s, nbr, start,
                       sorted, c.t_cols, c.t_vals, ustart, orow, ocol, oval);
    SGIPC_LAUNCH_CHECK();
    if (K > 0) {
        int blocks = div_up(9LL * K, 256);
        if (blocks > 65536) blocks = 65536;
        if (blocks < div_up(K, 256)) blocks = div_up(K, 256);
        hipLaunchKernelGGL(k_copy_back, dim3(blocks), dim3(256), 0, s, K, orow, ocol, oval, c.t_rows, c.t_cols,
                           c.t_vals);
        SGIPC_LAUNCH_CHECK();
    }
    c.K = K;
    c.have_bcoo = true;
    return K;
}

}  // namespace sgipc
