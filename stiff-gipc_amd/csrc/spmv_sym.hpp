// Device side of the symmetric block-3x3 SpMV over the upper BCOO.
//
// Reference: S/linear_system/utils/spmv.cu:8-133 (upper BCOO, head-segmented
// warp row sums plus FP64 atomics for the transposed half).  Here a group of
// G lanes owns one block row j and sums, in a fixed order, the transposed
// references A(i,j)^T x_i (i < j) and then the row's own upper blocks
// A(j,k) x_k, both listed in one per-row entry array.  The values are read in place from the converted BCOO
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

// Block k of the value array, column-major a[p + 3q] = A(p, q), transposed
// in registers when tr.  FP64: the BCOO itself (72 B per block, 8-byte
// aligned), read as four 16-byte and one 8-byte load.  FP32: the padded copy
// (48 B per block, columns in float4 lanes 0-2), three aligned 16-byte loads.
typedef double double2_a8 __attribute__((ext_vector_type(2), aligned(8)));
constexpr int kF32BlockStride = 12;  // floats per block in the FP32 copy
constexpr int kF32FullStride = 10;   // words per entry in the row-contiguous FP32 copy: 9 floats + the column (five aligned 8-byte pairs)

template <typename TV>
__device__ inline void load_block(const TV* __restrict__ vals, int k, bool tr, TV a[9]);

template <>
__device__ inline void load_block<double>(const double* __restrict__ vals, int k, bool tr, double a[9]) {
    const double* A = vals + 9 * (size_t)k;
    const double2_a8* A2 = reinterpret_cast<const double2_a8*>(A);
    double2_a8 u0 = A2[0], u1 = A2[1], u2 = A2[2], u3 = A2[3];
    a[0] = u0.x; a[1] = u0.y; a[2] = u1.x; a[3] = u1.y; a[4] = u2.x; a[5] = u2.y; a[6] = u3.x; a[7] = u3.y;
    a[8] = A[8];
    if (tr) {  // a[p + 3q] <-> a[q + 3p]
        double t;
        t = a[1]; a[1] = a[3]; a[3] = t;
        t = a[2]; a[2] = a[6]; a[6] = t;
        t = a[5]; a[5] = a[7]; a[7] = t;
    }
}

template <>
__device__ inline void load_block<float>(const float* __restrict__ vals, int k, bool tr, float a[9]) {
    const float4* A4 = reinterpret_cast<const float4*>(vals + kF32BlockStride * (size_t)k);
    float4 c0 = A4[0], c1 = A4[1], c2 = A4[2];
    a[0] = c0.x; a[1] = c0.y; a[2] = c0.z; a[3] = c1.x; a[4] = c1.y; a[5] = c1.z; a[6] = c2.x; a[7] = c2.y;
    a[8] = c2.z;
    if (tr) {
        float t;
        t = a[1]; a[1] = a[3]; a[3] = t;
        t = a[2]; a[2] = a[6]; a[6] = t;
        t = a[5]; a[5] = a[7]; a[7] = t;
    }
}

// one block-times-vector term: block k (transposed when tr) times x_j
template <typename TV, typename XF>
__device__ inline void block_term(const TV* __restrict__ vals, int k, int j, bool tr, const XF& xv, TV& s0, TV& s1,
                                  TV& s2) {
    TV a[9];
    load_block<TV>(vals, k, tr, a);
    TV x0 = (TV)xv(3 * (size_t)j), x1 = (TV)xv(3 * (size_t)j + 1), x2 = (TV)xv(3 * (size_t)j + 2);
    s0 += a[0] * x0 + a[3] * x1 + a[6] * x2;
    s1 += a[1] * x0 + a[4] * x1 + a[7] * x2;
    s2 += a[2] * x0 + a[5] * x1 + a[8] * x2;
}

// Partial row sums of this lane (before the group reduction).  Row `row`
// owns entries ent[eptr[row] .. eptr[row+1]) = {upper block index, column}:
// first its transposed references (column < row), then its own upper
// blocks.  Lane `sub` takes entries sub, sub+G, ...  xv(i) returns entry i
// of the vector being multiplied.
template <int G, typename TV, typename XF>
__device__ inline void sym_row_lane(int row, int sub, const int* __restrict__ eptr, const int2* __restrict__ ent,
                                    const TV* __restrict__ vals, const XF& xv, TV& s0, TV& s1, TV& s2) {
    const int e0 = eptr[row] + sub, end = eptr[row + 1];
    int e = e0;
    // two entries in flight per lane for 4-byte values; 8-byte values would
    // cost occupancy (registers) for no gain
    if (sizeof(TV) == 4 && e + G < end) {
        int2 q0 = ent[e], q1 = ent[e + G];
        TV a[9], b[9];
        load_block<TV>(vals, q0.x, q0.y < row, a);
        load_block<TV>(vals, q1.x, q1.y < row, b);
        TV x0 = (TV)xv(3 * (size_t)q0.y), x1 = (TV)xv(3 * (size_t)q0.y + 1), x2 = (TV)xv(3 * (size_t)q0.y + 2);
        TV y0 = (TV)xv(3 * (size_t)q1.y), y1 = (TV)xv(3 * (size_t)q1.y + 1), y2 = (TV)xv(3 * (size_t)q1.y + 2);
        s0 += a[0] * x0 + a[3] * x1 + a[6] * x2;
        s1 += a[1] * x0 + a[4] * x1 + a[7] * x2;
        s2 += a[2] * x0 + a[5] * x1 + a[8] * x2;
        s0 += b[0] * y0 + b[3] * y1 + b[6] * y2;
        s1 += b[1] * y0 + b[4] * y1 + b[7] * y2;
        s2 += b[2] * y0 + b[5] * y1 + b[8] * y2;
        e += 2 * G;
    }
    for (; e < end; e += G) {
        int2 q = ent[e];
        block_term<TV>(vals, q.x, q.y, q.y < row, xv, s0, s1, s2);
    }
}

// Row-contiguous FP32 copy of both triangles (mixed precision): entry e is
// its block (transposed where it is a transposed reference), 10 floats, read
// as two 16-byte loads at 8-byte alignment and one 8-byte load.
typedef float f4_a8 __attribute__((ext_vector_type(4), aligned(8)));
typedef float f3_a4 __attribute__((ext_vector_type(3), aligned(4)));

struct FullEntry {
    f4_a8 lo, hi;  // column-major elements 0-3, 4-7
    float a8;
    int col;       // block column (the entry's tenth word)
};
__device__ inline FullEntry load_full_entry(const float* __restrict__ vf, int e) {
    const float* A = vf + kF32FullStride * (size_t)e;
    FullEntry f;
    f.lo = *reinterpret_cast<const f4_a8*>(A);
    f.hi = *reinterpret_cast<const f4_a8*>(A + 4);
    const int2 t = *reinterpret_cast<const int2*>(A + 8);
    f.a8 = __int_as_float(t.x);
    f.col = t.y;
    return f;
}
template <typename V>
__device__ inline f3_a4 load_z3(const V* __restrict__ z, int c) {
    if constexpr (sizeof(V) == 4) {
        return *reinterpret_cast<const f3_a4*>(z + 3 * (size_t)c);
    } else {
        f3_a4 r;
        r.x = (float)z[3 * (size_t)c];
        r.y = (float)z[3 * (size_t)c + 1];
        r.z = (float)z[3 * (size_t)c + 2];
        return r;
    }
}
__device__ inline void full_madd(const FullEntry& f, const f3_a4& x, float& s0, float& s1, float& s2) {
    s0 += f.lo.x * x.x + f.lo.w * x.y + f.hi.z * x.z;
    s1 += f.lo.y * x.x + f.hi.x * x.y + f.hi.w * x.z;
    s2 += f.lo.z * x.x + f.hi.y * x.y + f.a8 * x.z;
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
