// Symmetric block-3x3 SpMV  y = a A x + b y  on the upper BCOO resident in
// the triplet view, with the transposed-reference structure built here.
//
// Reference: S/linear_system/utils/spmv.cu:8-133.  The transposed half is
// gathered per row from a list of references to upper blocks instead of
// scattered with atomics, so each row is written once and the sum order is
// fixed.
#include "spmv_sym.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

namespace sgipc {
namespace {

template <int G, typename TV, typename TX, typename TY>
__global__ void __launch_bounds__(256) k_spmv_sym(int nb, const int* __restrict__ eptr, const int2* __restrict__ ent,
                                                  const TV* __restrict__ vals, double a, const TX* __restrict__ x,
                                                  double b, TY* __restrict__ y) {
    constexpr int RPB = 256 / G;
    const int nchunks = (nb + RPB - 1) / RPB;
    ChunkWalk w = xcd_chunks(nchunks);
    for (int ch = w.first; ch < w.end; ch += w.step) {
        int row = ch * RPB + threadIdx.x / G, sub = threadIdx.x % G;
        TV s0 = 0, s1 = 0, s2 = 0;
        auto xv = [x](size_t i) { return x[i]; };
        if (row < nb) sym_row_lane<G, TV>(row, sub, eptr, ent, vals, xv, s0, s1, s2);
        group_sum3<G, TV>(s0, s1, s2);
        if (sub == 0 && row < nb) {
            size_t o = 3 * (size_t)row;
            if (b == 0.0) {
                y[o] = (TY)(a * s0); y[o + 1] = (TY)(a * s1); y[o + 2] = (TY)(a * s2);
            } else {
                y[o] = (TY)(a * s0 + b * (double)y[o]);
                y[o + 1] = (TY)(a * s1 + b * (double)y[o + 1]);
                y[o + 2] = (TY)(a * s2 + b * (double)y[o + 2]);
            }
        }
    }
}

// lower-reference counts per row: block (i, j), i < j, is referenced by row j
__global__ void k_low_count(int K, const int* __restrict__ r, const int* __restrict__ c, int* __restrict__ cnt) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    if (r[k] != c[k]) atomicAdd(cnt + c[k], 1);
}

// sort key of upper block k for the lower references: its column, or nb
// (after every column) for a diagonal block, which has none
__global__ void k_low_keys(int K, const int* __restrict__ r, const int* __restrict__ c, int nb,
                           unsigned* __restrict__ key) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    key[k] = r[k] == c[k] ? (unsigned)nb : (unsigned)c[k];
}

// references in (column, block index) order: block ids from the stable sort
__global__ void k_low_fill(int K, const int* __restrict__ nlow, const int* __restrict__ kk, const int* __restrict__ r,
                           int2* __restrict__ lref) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= K || q >= *nlow) return;
    const int k = kk[q];
    lref[q] = make_int2(k, r[k]);
}

// per-row entry list: transposed references, then the row's upper blocks
__global__ void k_entries(int nb, const int* __restrict__ lptr, const int2* __restrict__ lref,
                          const int* __restrict__ ustart, const int* __restrict__ ucol, int* __restrict__ eptr,
                          int2* __restrict__ ent) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > nb) return;
    int base = lptr[j] + ustart[j];
    eptr[j] = base;
    if (j == nb) return;
    int nl = lptr[j + 1] - lptr[j];
    for (int e = 0; e < nl; ++e) ent[base + e] = lref[lptr[j] + e];
    for (int k = ustart[j]; k < ustart[j + 1]; ++k) ent[base + nl + (k - ustart[j])] = make_int2(k, ucol[k]);
}

// FP32 copy of the BCOO values, 12 floats per block: column q in lanes
// 4q .. 4q+2, lane 4q+3 zero
__global__ void k_to_f32(long long K, const double* __restrict__ s, float* __restrict__ d) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < K * kF32BlockStride;
         i += (long long)gridDim.x * blockDim.x) {
        const long long k = i / kF32BlockStride;
        const int e = (int)(i % kF32BlockStride), q = e >> 2, p = e & 3;
        d[i] = (p < 3) ? (float)s[9 * k + p + 3 * q] : 0.0f;
    }
}

// entry-ordered FP32 copy: eight lanes per block row, lane j taking entries
// j, j + 8, ... of the row, so a row's lanes write adjacent 40-byte entries
// (nine floats and the column: five aligned 8-byte stores each); rows are
// walked in per-XCD ranges, as in the SpMV, so the transposed blocks a row
// gathers are shared in L2.
constexpr int kFullG = 8;
__global__ void __launch_bounds__(256) k_full_f32(int nb, const int* __restrict__ eptr, const int2* __restrict__ ent,
                                                  const double* __restrict__ vals, float* __restrict__ vf) {
    constexpr int RPB = 256 / kFullG;
    const ChunkWalk w = xcd_chunks((nb + RPB - 1) / RPB);
    for (int ch = w.first; ch < w.end; ch += w.step) {
        const int r = ch * RPB + threadIdx.x / kFullG, sub = threadIdx.x % kFullG;
        if (r >= nb) continue;
        const int end = eptr[r + 1];
        for (int e = eptr[r] + sub; e < end; e += kFullG) {
            const int2 q = ent[e];
            const double* A = vals + 9 * (size_t)q.x;
            double a[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) a[k] = A[k];
            float d[10];
            const bool tr = q.y < r;
#pragma unroll
            for (int pp = 0; pp < 3; ++pp)
#pragma unroll
                for (int qq = 0; qq < 3; ++qq) d[pp + 3 * qq] = (float)(tr ? a[qq + 3 * pp] : a[pp + 3 * qq]);
            d[9] = __int_as_float(q.y);
            float2* o = reinterpret_cast<float2*>(vf + kF32FullStride * (size_t)e);
#pragma unroll
            for (int k = 0; k < 5; ++k) o[k] = make_float2(d[2 * k], d[2 * k + 1]);
        }
    }
}

}  // namespace

// Build the transposed-reference structure for the BCOO now resident in the
// triplet view (rows/cols/vals [0, K)); uniq_start holds the row starts.
void bsr_build(Ctx& c) {
    if (!c.have_bcoo) throw state_error("bsr_build: convert first");
    hipStream_t s = c.stream;
    int nb = c.nb, K = c.K;
    SymBcoo& S = c.bsr;
    int* cnt = c.row_count.reserve((size_t)nb + 1);
    int* lptr = S.lptr.reserve((size_t)nb + 2);
    int* dint = c.dev_int.reserve(8);
    fill<int>(cnt, 0, nb, s);
    if (K > 0) {
        hipLaunchKernelGGL(k_low_count, dim3(div_up(K, 256)), dim3(256), 0, s, K, c.t_rows, c.t_cols, cnt);
        SGIPC_LAUNCH_CHECK();
    }
    exclusive_scan(c, cnt, lptr, nb, dint + 4, s);
    // the count of lower references (dint[4]) stays on the device: every
    // off-diagonal block has one, so K bounds it and sizes the buffers
    S.nb = nb;
    S.K = K;
    int2* lref = S.lref.reserve((size_t)K + 1);
    if (K > 0) {
        // the lower references of row j are the off-diagonal blocks of column
        // j, in block order (= row order): one stable radix sort by column
        int bc = 1;
        while ((1ll << bc) <= (long long)nb) ++bc;
        unsigned* key = S.lkey.reserve((size_t)K + 1);
        unsigned* key2 = S.lkey2.reserve((size_t)K + 1);
        int* kk = S.lidx.reserve((size_t)K + 1);
        hipLaunchKernelGGL(k_low_keys, dim3(div_up(K, 256)), dim3(256), 0, s, K, c.t_rows, c.t_cols, nb, key);
        SGIPC_LAUNCH_CHECK();
        if (c.sort_impl == SORT_OWN) {
            radix_sort_pairs(c, key, key2, nullptr, 0, kk, K, bc, s);
        } else {  // rocPRIM as the cross-check
            rocprim::counting_iterator<int> ids(0);
            size_t bytes = 0;
            SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, key, key2, ids, kk, (size_t)K, 0, bc, s));
            char* tmp = S.ltmp.reserve(bytes + 1);
            SGIPC_HIP(rocprim::radix_sort_pairs((void*)tmp, bytes, key, key2, ids, kk, (size_t)K, 0, bc, s));
        }
        hipLaunchKernelGGL(k_low_fill, dim3(div_up(K, 256)), dim3(256), 0, s, K, (const int*)(dint + 4),
                           (const int*)kk, c.t_rows, lref);
        SGIPC_LAUNCH_CHECK();
    }
    int* eptr = S.eptr.reserve((size_t)nb + 2);
    int2* ent = S.ent.reserve(2 * (size_t)K + 1);
    hipLaunchKernelGGL(k_entries, dim3(div_up(nb + 1, 256)), dim3(256), 0, s, nb, (const int*)lptr,
                       (const int2*)lref, (const int*)c.uniq_start.p, (const int*)c.t_cols, eptr, ent);
    SGIPC_LAUNCH_CHECK();
    S.valf_ok = false;
    S.full_ok = false;
    c.have_bsr = true;
}

// Row-contiguous FP32 copy of both triangles (mixed-precision SpMV)
const float* bsr_full_f32(Ctx& c) {
    if (!c.have_bsr) bsr_build(c);
    SymBcoo& S = c.bsr;
    const size_t ne = 2 * (size_t)S.K;  // bound on the entries (lower references + upper blocks)
    float* vf = S.valfull.reserve(kF32FullStride * ne + 4);
    if (!S.full_ok && S.nb > 0) {
        const int blocks = 8 * std::min(div_up(div_up(S.nb, 256 / kFullG), 8), 2048);
        hipLaunchKernelGGL(k_full_f32, dim3(blocks), dim3(256), 0, c.stream, S.nb, (const int*)S.eptr.p,
                           (const int2*)S.ent.p, (const double*)c.t_vals, vf);
        SGIPC_LAUNCH_CHECK();
    }
    S.full_ok = true;
    return vf;
}

// FP32 copy of the BCOO values for the mixed-precision solve
const float* bsr_values_f32(Ctx& c) {
    if (!c.have_bsr) bsr_build(c);
    SymBcoo& S = c.bsr;
    float* vf = S.valf.reserve((size_t)kF32BlockStride * S.K + 4);
    if (!S.valf_ok && S.K > 0) {
        long long n = (long long)kF32BlockStride * S.K;
        int g = (int)std::min<long long>(div_up(n, 256), 8192);
        hipLaunchKernelGGL(k_to_f32, dim3(g), dim3(256), 0, c.stream, (long long)S.K, (const double*)c.t_vals, vf);
        SGIPC_LAUNCH_CHECK();
    }
    S.valf_ok = true;
    return vf;
}

void spmv64(Ctx& c, double a, const double* x, double b, double* y) {
    if (!c.have_bsr) bsr_build(c);
    const SymBcoo& S = c.bsr;
    constexpr int G = 8;
    int nchunks = div_up((long long)S.nb, 256 / G);
    int blocks = std::max(8, (int)div_up(nchunks, 8) * 8);
    hipLaunchKernelGGL((k_spmv_sym<G, double, double, double>), dim3(blocks), dim3(256), 0, c.stream, S.nb, S.eptr.p,
                       S.ent.p, (const double*)c.t_vals, a, x, b, y);
    SGIPC_LAUNCH_CHECK();
}

}  // namespace sgipc
