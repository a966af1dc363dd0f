// Device side of the symmetric block-3x3 SpMV over the upper BCOO.
//
// Reference: S/linear_system/utils/spmv.cu:8-133 (upper BCOO, head-segmented
// warp row sums plus FP64 atomics for the transposed half).  Here a group of
// G lanes owns one block row j and sums, in a fixed order, the transposed
// references A(i,j)^T x_i (i < j) and then the row's own upper blocks
// A(j,k) x_k.  The values are read in place from the converted BCOO
// (9 doubles per block, column-major): no atomics, no second copy of the
// matrix, deterministic.
#pragma once
#include "context.hpp"

namespace sgipc {

// Rows are handled in chunks of (256 / G) rows.  The chunks are cut into 8
// contiguous ranges, one per XCD (consecutive workgroups are dispatched
// round-robin over the 8 XCDs), so the blocks a row references through the
// transpose -- blocks of nearby rows -- are usually still in that XCD's L2.
struct ChunkWalk {
    int first, step, end;
};
__device__ inline ChunkWalk xcd_chunks(int nchunks) {
    ChunkWalk w;
    int g = gridDim.x;
    if (g % 8 != 0) {
        w.first = blockIdx.x; w.step = g; w.end = nchunks;
        return w;
    }
    int per = g / 8, x = blockIdx.x % 8, loc = blockIdx.x / 8;
    int span = (nchunks + 7) / 8;
    w.first = x * span + loc;
    w.step = per;
    w.end = min((x + 1) * span, nchunks);
    return w;
}

template <typename TV>
__device__ inline void load_block(const TV* __restrict__ vals, int k, bool tr, TV a[9]) {
    const TV* A = vals + 9 * (size_t)k;
#pragma unroll
    for (int i = 0; i < 9; ++i) a[i] = A[i];
    if (tr) {  // transpose in registers: a[p + 3q] <-> a[q + 3p]
        TV t;
        t = a[1]; a[1] = a[3]; a[3] = t;
        t = a[2]; a[2] = a[6]; a[6] = t;
        t = a[5]; a[5] = a[7]; a[7] = t;
    }
}

// partial row sums of this lane (before the group reduction); xv(i) returns
// entry i of the vector being multiplied
template <int G, typename TV, typename XF>
__device__ inline void sym_row_lane(int row, int sub, const int* __restrict__ lptr, const int2* __restrict__ lref,
                                    const int* __restrict__ ustart, const int* __restrict__ ucol,
                                    const TV* __restrict__ vals, const XF& xv, TV& s0, TV& s1, TV& s2) {
    int l0 = lptr[row], nl = lptr[row + 1] - l0;
    int u0 = ustart[row], nu = ustart[row + 1] - u0;
    for (int e = sub; e < nl + nu; e += G) {
        int k, j;
        bool tr = e < nl;
        if (tr) {
            int2 q = lref[l0 + e];
            k = q.x; j = q.y;
        } else {
            k = u0 + (e - nl);
            j = ucol[k];
        }
        TV a[9];
        load_block<TV>(vals, k, tr, a);
        TV x0 = (TV)xv(3 * (size_t)j), x1 = (TV)xv(3 * (size_t)j + 1), x2 = (TV)xv(3 * (size_t)j + 2);
        s0 += a[0] * x0 + a[3] * x1 + a[6] * x2;
        s1 += a[1] * x0 + a[4] * x1 + a[7] * x2;
        s2 += a[2] * x0 + a[5] * x1 + a[8] * x2;
    }
}

template <int G, typename TV>
__device__ inline void group_sum3(TV& s0, TV& s1, TV& s2) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off, G);
        s1 += __shfl_xor(s1, off, G);
        s2 += __shfl_xor(s2, off, G);
    }
}

}  // namespace sgipc
