// Triplet -> upper block-3x3 BCOO (sorted by (row, col), duplicates summed).
//
// Reference: S/linear_system/utils/converter.cu:23-196 (radix sort of
// row<<32|col keys, run heads, exclusive scan, segmented block reduction with
// FP64 atomics).  Here: one stable device radix sort of (row << bc | col)
// keys over the triplet ids (bc = bits of the block row count), row bounds
// and unique-key heads read off the sorted keys, and a per-unique-block
// sequential sum in ascending triplet order -- deterministic, no
// floating-point atomics, and each 72-byte triplet value is read exactly
// once.
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
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

namespace sgipc {
namespace {

constexpr int kRowsPerBlock = 4;  // one wavefront per row

// Sort keys of the triplets [t0, t0 + n): row << bc | col.  A row or column
// outside [0, nbr) sets *err and takes the key nbr << bc, which sorts after
// every valid key (nbr < 2^bc).
__global__ void k_seg_keys(long long t0, long long n, const int* __restrict__ rows, const int* __restrict__ cols,
                           int nbr, int bc, unsigned long long* __restrict__ key, int* __restrict__ err) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = rows[t0 + i], c = cols[t0 + i];
    unsigned long long k;
    if ((unsigned)r >= (unsigned)nbr || (unsigned)c >= (unsigned)nbr) {
        atomicOr(err, 1);
        k = (unsigned long long)nbr << bc;
    } else {
        k = ((unsigned long long)r << bc) | (unsigned)c;
    }
    key[i] = k;
}

// First and one-past-last position of each valid row's run of sorted keys;
// rows without triplets keep beg = -1.
__global__ void k_row_bounds(long long n, const unsigned long long* __restrict__ key, int bc, int nbr,
                             int* __restrict__ beg, int* __restrict__ fin) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long r = key[i] >> bc;
    if (r >= (unsigned long long)nbr) return;
    if (i == 0 || (key[i - 1] >> bc) != r) beg[r] = (int)i;
    if (i == n - 1 || (key[i + 1] >> bc) != r) fin[r] = (int)(i + 1);
}

__global__ void k_row_counts(int nbr, const int* __restrict__ beg, const int* __restrict__ fin, int* __restrict__ cnt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nbr) cnt[r] = beg[r] < 0 ? 0 : fin[r] - beg[r];
}

// head[i] = 1 where sorted key i is valid and differs from key i - 1;
// head[n] = 0 (so the exclusive scan's entry n is the unique count)
__global__ void k_key_heads(long long n, const unsigned long long* __restrict__ key, int bc, int nbr,
                            int* __restrict__ head) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        head[n] = 0;
        return;
    }
    const unsigned long long k = key[i];
    head[i] = ((k >> bc) < (unsigned long long)nbr && (i == 0 || key[i - 1] != k)) ? 1 : 0;
}

// ustart[r] = unique blocks before row r (r in [0, nbr]); *total = all of them
__global__ void k_unique_starts(int nbr, const int* __restrict__ start, const int* __restrict__ hscan,
                                int* __restrict__ ustart, int* __restrict__ total) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nbr) return;
    const int u = hscan[start[r]];
    ustart[r] = u;
    if (r == nbr && total) *total = u;
}

// column and first sorted position of each unique block
__global__ void k_key_runs(long long n, const unsigned long long* __restrict__ key, int bc,
                           const int* __restrict__ head, const int* __restrict__ hscan, int* __restrict__ ucol,
                           int* __restrict__ urun) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !head[i]) return;
    const int u = hscan[i];
    ucol[u] = (int)(key[i] & ((1ull << bc) - 1ull));
    urun[u] = (int)i;
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

// Sort the triplets [t0, t1) by (row, col, triplet id) with one stable
// device radix sort: start (nbr + 1 row starts into sorted, positions
// relative to t0), sorted (triplet ids), ustart (nbr + 1 unique-block row
// starts; the total -> *total_dev), and, where ucol / urun are given, the
// column and first sorted position of each unique block.  Bad indices ->
// *err_dev (their triplets sort after every row and are left out).
static void sort_segment(Ctx& c, long long t0, long long t1, int nbr, int* start, int* sorted, int* ustart,
                         int* ucol, int* urun, int* err_dev, int* total_dev) {
    hipStream_t s = c.stream;
    const long long n = t1 - t0;
    int bc = 1;
    while ((1ll << bc) <= (long long)nbr) ++bc;  // nbr < 2^bc: the bad-index key fits
    unsigned long long* key = c.seg_key.reserve((size_t)n + 1);
    unsigned long long* key2 = c.seg_key2.reserve((size_t)n + 1);
    int* head = c.seg_head.reserve((size_t)n + 2);
    int* hscan = c.seg_hscan.reserve((size_t)n + 2);
    int* beg = c.seg_beg.reserve((size_t)nbr + 1);
    int* fin = c.seg_fin.reserve((size_t)nbr + 1);
    int* cnt = c.row_count.reserve((size_t)nbr + 1);
    const bool own = c.sort_impl == SORT_OWN;
    rocprim::counting_iterator<int> ids((int)t0);
    size_t bytes = 0;
    if (!own) {  // the library's sort and scan, or rocPRIM's as the cross-check
        size_t b1 = 0, b2 = 0;
        SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, b1, key, key2, ids, sorted, (size_t)n, 0, 2 * bc, s));
        SGIPC_HIP(rocprim::exclusive_scan(nullptr, b2, head, hscan, 0, (size_t)n + 1, rocprim::plus<int>(), s));
        bytes = std::max(b1, b2);
    }
    char* tmp = own ? nullptr : c.seg_tmp.reserve(bytes + 1);
    if (n > 0) {
        hipLaunchKernelGGL(k_seg_keys, dim3(div_up(n, 256)), dim3(256), 0, s, t0, n, c.t_rows, c.t_cols, nbr, bc, key,
                           err_dev);
        SGIPC_LAUNCH_CHECK();
        if (own) {
            radix_sort_pairs(c, key, key2, nullptr, (int)t0, sorted, n, 2 * bc, s);
        } else {
            size_t b = bytes;
            SGIPC_HIP(rocprim::radix_sort_pairs((void*)tmp, b, key, key2, ids, sorted, (size_t)n, 0, 2 * bc, s));
        }
    }
    fill<int>(beg, -1, nbr, s);
    if (n > 0) {
        hipLaunchKernelGGL(k_row_bounds, dim3(div_up(n, 256)), dim3(256), 0, s, n, (const unsigned long long*)key2, bc,
                           nbr, beg, fin);
        SGIPC_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_row_counts, dim3(div_up(nbr, 256)), dim3(256), 0, s, nbr, (const int*)beg, (const int*)fin,
                       cnt);
    SGIPC_LAUNCH_CHECK();
    exclusive_scan(c, cnt, start, nbr, nullptr, s);
    hipLaunchKernelGGL(k_key_heads, dim3(div_up(n + 1, 256)), dim3(256), 0, s, n, (const unsigned long long*)key2, bc,
                       nbr, head);
    SGIPC_LAUNCH_CHECK();
    if (own) {
        exclusive_scan(c, head, hscan, (int)(n + 1), nullptr, s);
    } else {
        size_t b = bytes;
        SGIPC_HIP(rocprim::exclusive_scan((void*)tmp, b, head, hscan, 0, (size_t)n + 1, rocprim::plus<int>(), s));
    }
    hipLaunchKernelGGL(k_unique_starts, dim3(div_up(nbr + 1, 256)), dim3(256), 0, s, nbr, (const int*)start,
                       (const int*)hscan, ustart, total_dev);
    SGIPC_LAUNCH_CHECK();
    if (ucol && n > 0) {
        hipLaunchKernelGGL(k_key_runs, dim3(div_up(n, 256)), dim3(256), 0, s, n, (const unsigned long long*)key2, bc,
                           (const int*)head, (const int*)hscan, ucol, urun);
        SGIPC_LAUNCH_CHECK();
    }
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
        int* start = c.row_start.reserve((size_t)nbr + 2);
        int* sorted = c.bcoo_row.reserve((size_t)nt + 1);
        sort_segment(c, 0, nt, nbr, start, sorted, ustart, nullptr, nullptr, dint + 2, dint);
        int hk[4] = {0, 0, 0, 0}, serr = 0;
        SGIPC_HIP(hipMemcpyAsync(hk, dint, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
        if (c.sort_err.p) SGIPC_HIP(hipMemcpyAsync(&serr, c.sort_err.p, sizeof(int), hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipStreamSynchronize(s));
        if (serr) throw std::runtime_error("convert: radix sort look-back failed");
        if (hk[2]) throw arg_error("convert: triplet row/col index out of range [0, n_block_rows)");
        const int K = hk[0];
        int* orow = c.merge_row.reserve((size_t)K + 1);
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
    int* dstart = c.row_start.reserve((size_t)nbr + 2);
    int* dsorted = c.bcoo_row.reserve((size_t)s0 + 1);
    int* dus = c.dyn_ustart.reserve((size_t)nbr + 2);
    int* ducol = c.dyn_ucol.reserve((size_t)s0 + 1);
    int* durun = c.dyn_urun.reserve((size_t)s0 + 1);
    sort_segment(c, 0, s0, nbr, dstart, dsorted, dus, ducol, durun, dint + 2, dint);
    // merged block count per row and K (into dint[1]) from the cached plan,
    // before the check of that plan is read, so one host read serves both
    // (redone below when the plan had to be rebuilt)
    auto merge_count = [&]() {
        int* mcnt = c.row_count.get();  // the sort's row counts are no longer needed
        hipLaunchKernelGGL(k_merge_count, dim3(div_up(nbr, 256)), dim3(256), 0, s, nbr, (const int*)P.ustart.p,
                           (const int*)P.ucol.p, (const int*)dus, (const int*)ducol, mcnt);
        SGIPC_LAUNCH_CHECK();
        exclusive_scan(c, mcnt, ustart, nbr, dint + 1, s);
    };
    if (candidate) merge_count();
    int hk[8] = {0}, serr = 0;
    SGIPC_HIP(hipMemcpyAsync(hk, dint, sizeof(int) * 8, hipMemcpyDeviceToHost, s));
    if (c.sort_err.p) SGIPC_HIP(hipMemcpyAsync(&serr, c.sort_err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    if (serr) throw std::runtime_error("convert: radix sort look-back failed");
    if (hk[2]) throw arg_error("convert: triplet row/col index out of range [0, n_block_rows)");
    int K = hk[1];
    if (!candidate || hk[5]) {
        // (re)build the plan of the static segment
        P.valid = false;
        int* pstart = P.start.reserve((size_t)nbr + 2);
        int* psorted = P.sorted.reserve((size_t)ns + 1);
        int* pus = P.ustart.reserve((size_t)nbr + 2);
        int* pucol = P.ucol.reserve((size_t)ns + 1);
        int* purun = P.urun.reserve((size_t)ns + 1);
        fill<int>(dint + 6, 0, 2, s);
        sort_segment(c, s0, nt, nbr, pstart, psorted, pus, pucol, purun, dint + 6, dint + 7);
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
        merge_count();
        SGIPC_HIP(hipMemcpyAsync(&K, dint + 1, sizeof(int), hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipStreamSynchronize(s));
    }
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
