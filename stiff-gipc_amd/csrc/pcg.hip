// Preconditioned conjugate gradients on the converted block system.
//
// Reference: S/linear_system/solver/pcg_solver.cu:158-267 (PCGSolver::pcg),
// dot products :8-150 (two-stage warp reduction + blocking cudaMemcpy per
// dot), vector updates :105-122; config pcg_solver.h:6-16.
//
// Here every scalar (rz, alpha, beta, the stop decision) lives in device
// memory: the dot products finish in the last workgroup of the producing
// kernel (fixed-order, deterministic), alpha/beta are computed on the GPU,
// and every per-iteration kernel exits at once when the device `done` flag
// is set.  The host only polls that flag every few iterations, so the
// iteration count and the iterate are exactly those of the reference loop
// (stop test on the stale rz after the x/r update).
#include "context.hpp"
#include <cmath>

namespace sgipc {
namespace {

enum { S_RZ0 = 0, S_BB, S_TOL, S_NUM };
enum { C_STOP = 0, C_NUM };  // C_STOP: iteration at which the loop stopped (0 = running)

constexpr int kThreads = 256;
constexpr int kRedBlocks = 512;   // producers of dot partials use at most this many blocks
constexpr int kPart = 1024;       // partial slots per dot

__device__ inline double block_sum_all(double v, double* sh) {
    // result broadcast to every thread; fixed reduction tree
    v = wave_sum(v);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
    return s;
}

// sum of np partials, computed redundantly (same order) by every block
__device__ inline double sum_partials(const double* __restrict__ part, int np, double* sh) {
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[i];
    return block_sum_all(v, sh);
}

__device__ inline void write_partial(double v, double* __restrict__ part, double* sh) {
    double s = block_sum_all(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// partial dot products of a.b, one per block
__global__ void __launch_bounds__(kThreads) k_dot_part(long long N, const double* __restrict__ a,
                                                       const double* __restrict__ b, double* __restrict__ part,
                                                       const int* __restrict__ ctrl) {
    if (ctrl && ctrl[C_STOP]) return;
    __shared__ double sh[kThreads / 64];
    double v = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
        v += a[i] * b[i];
    write_partial(v, part, sh);
}

// Ap = A p (full block-CSR) with per-block partials of p.Ap
template <int G>
__global__ void __launch_bounds__(kThreads) k_spmv_pap(int nb, const int* __restrict__ rowptr, const int* __restrict__ col,
                                                       const double* __restrict__ val, int nnzb,
                                                       const double* __restrict__ p, double* __restrict__ Ap,
                                                       double* __restrict__ part, const int* __restrict__ ctrl) {
    if (ctrl[C_STOP]) return;
    __shared__ double sh[kThreads / 64];
    double contrib = 0.0;
    for (int base = blockIdx.x * (kThreads / G); base < nb; base += gridDim.x * (kThreads / G)) {
        int row = base + threadIdx.x / G, sub = threadIdx.x % G;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        if (row < nb) {
            int e1 = rowptr[row + 1];
            for (int e = rowptr[row] + sub; e < e1; e += G) {
                int j = col[e];
                double x0 = p[3 * (size_t)j], x1 = p[3 * (size_t)j + 1], x2 = p[3 * (size_t)j + 2];
                s0 += val[(size_t)0 * nnzb + e] * x0 + val[(size_t)3 * nnzb + e] * x1 + val[(size_t)6 * nnzb + e] * x2;
                s1 += val[(size_t)1 * nnzb + e] * x0 + val[(size_t)4 * nnzb + e] * x1 + val[(size_t)7 * nnzb + e] * x2;
                s2 += val[(size_t)2 * nnzb + e] * x0 + val[(size_t)5 * nnzb + e] * x1 + val[(size_t)8 * nnzb + e] * x2;
            }
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) {
            s0 += __shfl_xor(s0, off, G);
            s1 += __shfl_xor(s1, off, G);
            s2 += __shfl_xor(s2, off, G);
        }
        if (sub == 0 && row < nb) {
            size_t o = 3 * (size_t)row;
            Ap[o] = s0; Ap[o + 1] = s1; Ap[o + 2] = s2;
            contrib += p[o] * s0 + p[o + 1] * s1 + p[o + 2] * s2;
        }
    }
    write_partial(contrib, part, sh);
}

// alpha = rz / p.Ap ; x += alpha p ; r -= alpha Ap ; partials of r.r.
// Reference stop test (|rz| <= tol rz0 on the pre-update rz) decided here.
__global__ void __launch_bounds__(kThreads) k_update_xr(long long N, double* __restrict__ x, double* __restrict__ r,
                                                        const double* __restrict__ p, const double* __restrict__ Ap,
                                                        const double* __restrict__ part_pap, int np_pap,
                                                        const double* __restrict__ part_rz, int np_rz,
                                                        double* __restrict__ part_rr, double* __restrict__ sc,
                                                        int* __restrict__ ctrl, int iter, int rule) {
    int stop = ctrl[C_STOP];
    if (stop && stop < iter) return;
    __shared__ double sh[kThreads / 64];
    double rz = sum_partials(part_rz, np_rz, sh);
    double pap = sum_partials(part_pap, np_pap, sh);
    double al = rz / pap;
    double v = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        x[i] += al * p[i];
        double ri = r[i] - al * Ap[i];
        r[i] = ri;
        v += ri * ri;
    }
    write_partial(v, part_rr, sh);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double rz0 = (iter == 1) ? rz : sc[S_RZ0];
        if (iter == 1) sc[S_RZ0] = rz;
        if (rule == 0 && fabs(rz) <= sc[S_TOL] * rz0) ctrl[C_STOP] = iter;
    }
}

// true-relative stop test ||r|| <= tol ||b|| (one block)
__global__ void __launch_bounds__(kThreads) k_check_rr(const double* __restrict__ part_rr, int np, const double* __restrict__ sc,
                                                       int* __restrict__ ctrl, int iter) {
    if (ctrl[C_STOP]) return;
    __shared__ double sh[kThreads / 64];
    double rr = sum_partials(part_rr, np, sh);
    if (threadIdx.x == 0 && sqrt(rr) <= sc[S_TOL] * sqrt(sc[S_BB])) ctrl[C_STOP] = iter;
}

// beta = rz_new / rz_old ; p = z + beta p
__global__ void __launch_bounds__(kThreads) k_update_p(long long N, double* __restrict__ p, const double* __restrict__ z,
                                                       const double* __restrict__ part_new, const double* __restrict__ part_old,
                                                       int np, const int* __restrict__ ctrl) {
    if (ctrl[C_STOP]) return;
    __shared__ double sh[kThreads / 64];
    double rn = sum_partials(part_new, np, sh);
    double ro = sum_partials(part_old, np, sh);
    double be = rn / ro;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
        p[i] = z[i] + be * p[i];
}

// scalar = sum of partials into sc[idx] (one block)
__global__ void __launch_bounds__(kThreads) k_store_sum(const double* __restrict__ part, int np, double* __restrict__ sc,
                                                        int idx) {
    __shared__ double sh[kThreads / 64];
    double s = sum_partials(part, np, sh);
    if (threadIdx.x == 0) sc[idx] = s;
}

__global__ void k_copy_skip(long long N, double* __restrict__ dst, const double* __restrict__ src,
                            const int* __restrict__ ctrl) {
    if (ctrl && ctrl[C_STOP]) return;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_diag_inv(int K, const int* __restrict__ rows, const int* __restrict__ cols,
                           const double* __restrict__ vals, double* __restrict__ Dinv) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K || rows[k] != cols[k]) return;
    const double* M = vals + 9 * (size_t)k;  // column-major
    double a = M[0], b = M[3], c = M[6], d = M[1], e = M[4], f = M[7], g = M[2], h = M[5], i = M[8];
    double A = e * i - f * h, Bq = -(d * i - f * g), C = d * h - e * g;
    double id = 1.0 / (a * A + b * Bq + c * C);
    double* R = Dinv + 9 * (size_t)rows[k];
    R[0] = A * id; R[3] = -(b * i - c * h) * id; R[6] = (b * f - c * e) * id;
    R[1] = Bq * id; R[4] = (a * i - c * g) * id; R[7] = -(a * f - c * d) * id;
    R[2] = C * id; R[5] = -(a * h - b * g) * id; R[8] = (a * e - b * d) * id;
}

__global__ void k_diag_apply(int nb, const double* __restrict__ Dinv, const double* __restrict__ r,
                             double* __restrict__ z, const int* __restrict__ ctrl) {
    if (ctrl && ctrl[C_STOP]) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const double* D = Dinv + 9 * (size_t)i;
    double r0 = r[3 * (size_t)i], r1 = r[3 * (size_t)i + 1], r2 = r[3 * (size_t)i + 2];
    z[3 * (size_t)i] = D[0] * r0 + D[3] * r1 + D[6] * r2;
    z[3 * (size_t)i + 1] = D[1] * r0 + D[4] * r1 + D[7] * r2;
    z[3 * (size_t)i + 2] = D[2] * r0 + D[5] * r1 + D[8] * r2;
}

int grid_for(long long N) {
    int g = div_up(N, kThreads);
    return g < 1 ? 1 : (g > kRedBlocks ? kRedBlocks : g);
}

}  // namespace

void diag_build(Ctx& c) {
    if (!c.have_bcoo) throw state_error("diag preconditioner: convert first");
    double* D = c.pcg.Dinv.reserve(9 * (size_t)c.nb + 1);
    fill<double>(D, 0.0, 9LL * c.nb, c.stream);
    if (c.K > 0) {
        hipLaunchKernelGGL(k_diag_inv, dim3(div_up(c.K, 256)), dim3(256), 0, c.stream, c.K, c.t_rows, c.t_cols, c.t_vals, D);
        SGIPC_LAUNCH_CHECK();
    }
}

// z = M^-1 r with the configured preconditioner; `ctrl` (device) lets the
// PCG loop skip work after convergence (nullptr outside PCG).
static void precond_apply_ctrl(Ctx& c, const double* r, double* z, const int* ctrl) {
    long long N = 3LL * c.nb;
    hipStream_t s = c.stream;
    if (c.precond_kind == 1) {
        hipLaunchKernelGGL(k_diag_apply, dim3(div_up(c.nb, 256)), dim3(256), 0, s, c.nb, c.pcg.Dinv.p, r, z, ctrl);
        SGIPC_LAUNCH_CHECK();
        return;
    }
    if (c.precond_kind == 2 || c.precond_kind == 3) {
        long long off = 3LL * c.sub_offset, len = 3LL * c.mas.n;
        if (off > 0 || off + len < N) {  // identity outside the MAS subsystem range
            hipLaunchKernelGGL(k_copy_skip, dim3(grid_for(N)), dim3(kThreads), 0, s, N, z, r, ctrl);
            SGIPC_LAUNCH_CHECK();
        }
        if (c.precond_kind == 2) mas_apply64(c, r + off, z + off);
        else mas_apply32(c, r + off, z + off);
        return;
    }
    hipLaunchKernelGGL(k_copy_skip, dim3(grid_for(N)), dim3(kThreads), 0, s, N, z, r, ctrl);
    SGIPC_LAUNCH_CHECK();
}

void precond_apply(Ctx& c, const double* r, double* z) { precond_apply_ctrl(c, r, z, nullptr); }

void pcg_solve(Ctx& c, const double* b, double* x, const SolveOpts& o, SolveStats& st) {
    if (!c.have_bcoo) throw state_error("solve: convert first");
    if (!c.have_bsr) bsr_build(c);
    if (o.precision != 0) throw arg_error("solve: mixed precision is not available in this build");
    hipStream_t s = c.stream;
    const long long N = 3LL * c.nb;
    PcgWork& w = c.pcg;
    double* r = w.r.reserve(N + 1);
    double* z = w.z.reserve(N + 1);
    double* p = w.p.reserve(N + 1);
    double* Ap = w.Ap.reserve(N + 1);
    double* parts = w.partials.reserve(5 * kPart);
    double* part_rz[2] = {parts, parts + kPart};
    double* part_pap = parts + 2 * kPart;
    double* part_rr = parts + 3 * kPart;
    double* part_bb = parts + 4 * kPart;
    double* sc = w.scal.reserve(S_NUM + 8);
    int* ctrl = w.ctrl.reserve(C_NUM + 8);
    fill<int>(ctrl, 0, C_NUM, s);
    double host_sc[S_NUM] = {0};
    host_sc[S_TOL] = o.tol;
    SGIPC_HIP(hipMemcpyAsync(sc, host_sc, sizeof(host_sc), hipMemcpyHostToDevice, s));
    const int max_iter = o.max_iter > 0 ? o.max_iter : (int)(o.max_iter_ratio * (double)N);
    const int gN = grid_for(N);
    const Bsr& B = c.bsr;
    constexpr int G = 8;
    const int gS = std::max(1, std::min(kPart, div_up((long long)B.nb * G, kThreads)));
    fill<double>(x, 0.0, N, s);
    hipLaunchKernelGGL(k_copy_skip, dim3(gN), dim3(kThreads), 0, s, N, r, b, (const int*)nullptr);
    SGIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dot_part, dim3(gN), dim3(kThreads), 0, s, N, b, b, part_bb, (const int*)nullptr);
    SGIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_store_sum, dim3(1), dim3(kThreads), 0, s, part_bb, gN, sc, (int)S_BB);
    SGIPC_LAUNCH_CHECK();
    precond_apply_ctrl(c, r, z, nullptr);
    hipLaunchKernelGGL(k_dot_part, dim3(gN), dim3(kThreads), 0, s, N, r, z, part_rz[0], (const int*)nullptr);
    SGIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_store_sum, dim3(1), dim3(kThreads), 0, s, part_rz[0], gN, sc, (int)S_RZ0);
    SGIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_copy_skip, dim3(gN), dim3(kThreads), 0, s, N, p, z, (const int*)nullptr);
    SGIPC_LAUNCH_CHECK();
    double rz0 = 0.0;
    SGIPC_HIP(hipMemcpyAsync(&rz0, sc + S_RZ0, sizeof(double), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    st.rz0 = rz0;
    int stop_iter = 0;
    if (rz0 == 0.0) {
        // zero right-hand side: x = 0 is exact (the reference would divide 0/0)
        st.iterations = 0;
        st.converged = 1;
    } else {
        const int check_every = o.check_every > 0 ? o.check_every : 8;
        int k = 1;
        for (; k < max_iter; ++k) {
            prof_begin(c, PROF_SPMV);
            hipLaunchKernelGGL(k_spmv_pap<G>, dim3(gS), dim3(kThreads), 0, s, B.nb, B.rowptr.p, B.col.p, B.val.p,
                               B.nnzb, p, Ap, part_pap, ctrl);
            SGIPC_LAUNCH_CHECK();
            prof_end(c, PROF_SPMV);
            hipLaunchKernelGGL(k_update_xr, dim3(gN), dim3(kThreads), 0, s, N, x, r, p, Ap, part_pap, gS,
                               part_rz[(k - 1) & 1], gN, part_rr, sc, ctrl, k, o.stop_rule);
            SGIPC_LAUNCH_CHECK();
            if (o.stop_rule == 1) {
                hipLaunchKernelGGL(k_check_rr, dim3(1), dim3(kThreads), 0, s, part_rr, gN, sc, ctrl, k);
                SGIPC_LAUNCH_CHECK();
            }
            precond_apply_ctrl(c, r, z, ctrl);
            hipLaunchKernelGGL(k_dot_part, dim3(gN), dim3(kThreads), 0, s, N, r, z, part_rz[k & 1], ctrl);
            SGIPC_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_update_p, dim3(gN), dim3(kThreads), 0, s, N, p, z, part_rz[k & 1],
                               part_rz[(k - 1) & 1], gN, ctrl);
            SGIPC_LAUNCH_CHECK();
            if ((k % check_every) == 0 || k + 1 >= max_iter) {
                SGIPC_HIP(hipMemcpyAsync(&stop_iter, ctrl + C_STOP, sizeof(int), hipMemcpyDeviceToHost, s));
                SGIPC_HIP(hipStreamSynchronize(s));
                if (stop_iter) break;
            }
        }
        SGIPC_HIP(hipMemcpyAsync(&stop_iter, ctrl + C_STOP, sizeof(int), hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipStreamSynchronize(s));
        st.converged = stop_iter ? 1 : 0;
        st.iterations = stop_iter ? stop_iter : max_iter;
    }
    // true residual ||b - A x|| / ||b||, FP64 (tmp = b - A x)
    double* tmp = w.xw.reserve(N + 1);
    hipLaunchKernelGGL(k_copy_skip, dim3(gN), dim3(kThreads), 0, s, N, tmp, b, (const int*)nullptr);
    SGIPC_LAUNCH_CHECK();
    spmv64(c, -1.0, x, 1.0, tmp);
    hipLaunchKernelGGL(k_dot_part, dim3(gN), dim3(kThreads), 0, s, N, tmp, tmp, part_rr, (const int*)nullptr);
    SGIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_store_sum, dim3(1), dim3(kThreads), 0, s, part_rr, gN, sc, (int)S_RZ0);
    SGIPC_LAUNCH_CHECK();
    double rr = 0.0, bb = 0.0;
    SGIPC_HIP(hipMemcpyAsync(&rr, sc + S_RZ0, sizeof(double), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipMemcpyAsync(&bb, sc + S_BB, sizeof(double), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    st.rel_residual = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
}

}  // namespace sgipc
