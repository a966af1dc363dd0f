// Device-side pieces shared by the PCG kernels and the MAS kernels that run
// inside a PCG iteration: the control words, the device scalars and the
// fixed-order workgroup reductions used for every dot product.
#pragma once
#include "common.hpp"

namespace sgipc {

// device scalars: rz0, b.b, tolerance, progress of the stop test (the
// ratio compared against the tolerance at the last test)
enum { S_RZ0 = 0, S_BB, S_TOL, S_PROG, S_NUM };
// control words: C_STOP = iteration at which the loop stopped (0 = running),
// C_ITER = current iteration (from 1), C_MAXIT = iteration budget,
// C_NOCONV = stopped by the budget
enum { C_STOP = 0, C_ITER, C_MAXIT, C_NOCONV, C_NUM };

constexpr int kPart = 1024;  // partial slots per dot product

// workgroup sum broadcast to every thread; fixed reduction tree
__device__ inline double block_sum_all(double v, double* sh) {
    v = wave_sum(v);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
    return s;
}

// two workgroup sums at once (each with the tree of block_sum_all, so the
// same values), sharing the barriers; sh holds 2 * (waves) doubles
__device__ inline void block_sum_all2(double& a, double& b, double* sh) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    __syncthreads();
    if (lane == 0) { sh[w] = a; sh[nw + w] = b; }
    __syncthreads();
    double sa = 0.0, sb = 0.0;
    for (int i = 0; i < nw; ++i) { sa += sh[i]; sb += sh[nw + i]; }
    a = sa;
    b = sb;
}

// sums of two partial sets, as two sum_partials calls would give them
__device__ inline void sum_partials2(const double* __restrict__ pa, int na, const double* __restrict__ pb, int nb,
                                     double* sh, double& a, double& b) {
    a = 0.0;
    b = 0.0;
    for (int i = threadIdx.x; i < na; i += blockDim.x) a += pa[i];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) b += pb[i];
    block_sum_all2(a, b, sh);
}

// sum of np partials, computed redundantly (same order) by every workgroup
__device__ inline double sum_partials(const double* __restrict__ part, int np, double* sh) {
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[i];
    return block_sum_all(v, sh);
}

// one partial per workgroup, at part[blockIdx.x]
__device__ inline void write_partial(double v, double* __restrict__ part, double* sh) {
    double s = block_sum_all(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

constexpr int C_PUB = C_NUM + 2;  // publications of the progress words (pinned polling)
constexpr int kLevelBatch = 8;    // MAS levels whose loads are in flight together in the collect

// Stop tests, evaluated identically by every block of the kernel that
// follows the x/r update (so no block depends on another's flag write):
// rule 0 (reference): |rz_prev| <= tol * rz0 on the pre-update rz;
// rule 1: ||r|| <= tol * ||b|| on the updated residual.  rule < 0: none.
// This thread's share of the stop test's partial sum (the strided part of
// sum_partials), so a kernel can issue its loads early and reduce late.
__device__ inline double stop_share(int rule, const double* __restrict__ part_rz_prev, int np_rz,
                                    const double* __restrict__ part_rr, int np_rr) {
    double v = 0.0;
    if (rule == 0) {
        for (int i = threadIdx.x; i < np_rz; i += blockDim.x) v += part_rz_prev[i];
    } else if (rule > 0) {
        for (int i = threadIdx.x; i < np_rr; i += blockDim.x) v += part_rr[i];
    }
    return v;
}

// Stop test from the shares: rule 0 |r.z| <= tol * r0.z0 (previous r.z),
// rule 1 ||r|| <= tol * ||b||; block 0 records the progress and the stop.
__device__ inline bool stop_test_share(int rule, double share, double* __restrict__ sc, int* __restrict__ ctrl,
                                       double* sh) {
    if (rule < 0) return false;
    const double sum = block_sum_all(share, sh);
    bool stop;
    double q;  // progress: the ratio compared with the tolerance
    if (rule == 0) {
        stop = fabs(sum) <= sc[S_TOL] * sc[S_RZ0];
        q = fabs(sum) / sc[S_RZ0];
    } else {
        stop = sqrt(sum) <= sc[S_TOL] * sqrt(sc[S_BB]);
        q = sqrt(sum) / sqrt(sc[S_BB]);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc[S_PROG] = q;
        if (stop) ctrl[C_STOP] = ctrl[C_ITER] - 1;
    }
    return stop;
}

__device__ inline bool stop_test(int rule, const double* __restrict__ part_rz_prev, int np_rz,
                                 const double* __restrict__ part_rr, int np_rr, double* __restrict__ sc,
                                 int* __restrict__ ctrl, double* sh) {
    return stop_test_share(rule, stop_share(rule, part_rz_prev, np_rz, part_rr, np_rr), sc, ctrl, sh);
}

// Progress words for the host's pinned-memory polling: the stop and
// iteration words and the progress ratio, then (after a system-scope fence)
// the publication count the host waits on.  One thread.
__device__ inline void publish_progress(int* __restrict__ ctrl, const double* __restrict__ sc, volatile int* hp,
                                        volatile double* hq) {
    const int pub = ctrl[C_PUB] + 1;
    ctrl[C_PUB] = pub;
    hq[0] = sc[S_PROG];
    hp[0] = ctrl[C_STOP];
    hp[1] = ctrl[C_ITER];
    __threadfence_system();
    hp[2] = pub;
}

}  // namespace sgipc
