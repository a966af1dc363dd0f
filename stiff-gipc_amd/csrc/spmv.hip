// Symmetric block-3x3 SpMV  y = a A x + b y  on the full block-CSR built from
// the upper BCOO (both triangles stored, values SoA by component).
//
// Reference: S/linear_system/utils/spmv.cu:8-133 (upper BCOO, warp
// head-segmented row sums, FP64 atomics for the transposed half).  Here each
// row is owned by a group of G lanes: a gather over the row's blocks, a
// shuffle reduction inside the group, one store per row -- no atomics,
// deterministic, coalesced SoA value streams.
#include "context.hpp"

namespace sgipc {

template <int G, typename TV, typename TX, typename TY>
__global__ void __launch_bounds__(256) k_spmv_bsr(int nb, const int* __restrict__ rowptr,
                                                  const int* __restrict__ col, const TV* __restrict__ val,
                                                  int nnzb, double a, const TX* __restrict__ x, double b,
                                                  TY* __restrict__ y) {
    int gid = (blockIdx.x * blockDim.x + threadIdx.x);
    int row = gid / G, sub = gid % G;
    if (row >= nb) return;  // G divides 64, so whole groups leave together
    int e0 = rowptr[row], e1 = rowptr[row + 1];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int e = e0 + sub; e < e1; e += G) {
        int j = col[e];
        double x0 = (double)x[3 * (size_t)j], x1 = (double)x[3 * (size_t)j + 1], x2 = (double)x[3 * (size_t)j + 2];
        // val[(p + 3q) * nnzb + e] = A(p,q)
        s0 += (double)val[(size_t)0 * nnzb + e] * x0 + (double)val[(size_t)3 * nnzb + e] * x1 +
              (double)val[(size_t)6 * nnzb + e] * x2;
        s1 += (double)val[(size_t)1 * nnzb + e] * x0 + (double)val[(size_t)4 * nnzb + e] * x1 +
              (double)val[(size_t)7 * nnzb + e] * x2;
        s2 += (double)val[(size_t)2 * nnzb + e] * x0 + (double)val[(size_t)5 * nnzb + e] * x1 +
              (double)val[(size_t)8 * nnzb + e] * x2;
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off, G);
        s1 += __shfl_xor(s1, off, G);
        s2 += __shfl_xor(s2, off, G);
    }
    if (sub == 0) {
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

void spmv64(Ctx& c, double a, const double* x, double b, double* y) {
    if (!c.have_bsr) bsr_build(c);
    const Bsr& B = c.bsr;
    constexpr int G = 8;
    int threads = 256;
    int blocks = div_up((long long)B.nb * G, threads);
    hipLaunchKernelGGL((k_spmv_bsr<G, double, double, double>), dim3(blocks), dim3(threads), 0, c.stream, B.nb,
                       B.rowptr.p, B.col.p, B.val.p, B.nnzb, a, x, b, y);
    SGIPC_LAUNCH_CHECK();
}

}  // namespace sgipc
