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

enum { S_RZ = 0, S_RZ0, S_ALPHA, S_BETA, S_PAP, S_RR, S_BB, S_TOL, S_NUM };
enum { C_DONE = 0, C_ITERS, C_T0, C_T1, C_T2, C_STOP_AFTER, C_NUM };

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 2048;

__device__ inline double block_sum(double v, double* sh) {
    v = wave_sum(v);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
    return s;  // valid in thread 0
}

// last-workgroup finalisation: returns true in thread 0 of the last block,
// with the fixed-order total of partials[] in *total.
__device__ inline bool last_block(double my_partial, double* partials, int* ticket, double* total) {
    __shared__ bool is_last;
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = my_partial;
        __threadfence();
        int t = atomicAdd(ticket, 1);
        is_last = (t == (int)gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < (int)gridDim.x; ++i) s += ((volatile double*)partials)[i];
        *total = s;
        *ticket = 0;
    }
    return threadIdx.x == 0;
}

enum { OP_INIT = 0, OP_BETA = 1, OP_BNORM = 2 };

// dot(a, b) with a finishing operation on the device scalars
template <int OP>
__global__ void __launch_bounds__(kThreads) k_dot(long long N, const double* __restrict__ a, const double* __restrict__ b,
                                                  double* __restrict__ partials, double* __restrict__ sc,
                                                  int* __restrict__ ctrl) {
    if (OP == OP_BETA && ctrl[C_DONE]) return;
    __shared__ double sh[kThreads / 64];
    double v = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
        v += a[i] * b[i];
    double bs = block_sum(v, sh);
    double tot;
    if (last_block(bs, partials, ctrl + C_T1, &tot)) {
        if (OP == OP_INIT) {
            sc[S_RZ] = tot;
            sc[S_RZ0] = tot;
            if (tot == 0.0) { ctrl[C_DONE] = 1; ctrl[C_ITERS] = 0; }
        } else if (OP == OP_BETA) {
            sc[S_BETA] = tot / sc[S_RZ];
            sc[S_RZ] = tot;
        } else {
            sc[S_BB] = tot;
        }
    }
}

// Ap = A p (full block-CSR) and alpha = rz / (p.Ap); reference stop test.
template <int G>
__global__ void __launch_bounds__(kThreads) k_spmv_pap(int nb, const int* __restrict__ rowptr, const int* __restrict__ col,
                                                       const double* __restrict__ val, int nnzb,
                                                       const double* __restrict__ p, double* __restrict__ Ap,
                                                       double* __restrict__ partials, double* __restrict__ sc,
                                                       int* __restrict__ ctrl, int iter, int rule) {
    if (ctrl[C_DONE]) return;
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
    double bs = block_sum(contrib, sh);
    double tot;
    if (last_block(bs, partials, ctrl + C_T0, &tot)) {
        sc[S_PAP] = tot;
        sc[S_ALPHA] = sc[S_RZ] / tot;
        if (rule == 0 && fabs(sc[S_RZ]) <= sc[S_TOL] * sc[S_RZ0]) {
            ctrl[C_STOP_AFTER] = 1;
            ctrl[C_ITERS] = iter;
        }
    }
}

// x += alpha p ; r -= alpha Ap ; optional true-relative stop test on ||r||
__global__ void __launch_bounds__(kThreads) k_update_xr(long long N, double* __restrict__ x, double* __restrict__ r,
                                                        const double* __restrict__ p, const double* __restrict__ Ap,
                                                        double* __restrict__ partials, double* __restrict__ sc,
                                                        int* __restrict__ ctrl, int iter, int rule) {
    if (ctrl[C_DONE]) return;
    __shared__ double sh[kThreads / 64];
    double al = sc[S_ALPHA];
    double v = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        x[i] += al * p[i];
        double ri = r[i] - al * Ap[i];
        r[i] = ri;
        v += ri * ri;
    }
    double bs = block_sum(v, sh);
    double tot;
    if (last_block(bs, partials, ctrl + C_T2, &tot)) {
        sc[S_RR] = tot;
        if (rule == 0) {
            if (ctrl[C_STOP_AFTER]) ctrl[C_DONE] = 1;
        } else if (sqrt(tot) <= sc[S_TOL] * sqrt(sc[S_BB])) {
            ctrl[C_DONE] = 1;
            ctrl[C_ITERS] = iter;
        }
    }
}

__global__ void k_update_p(long long N, double* __restrict__ p, const double* __restrict__ z,
                           const double* __restrict__ sc, const int* __restrict__ ctrl) {
    if (ctrl[C_DONE]) return;
    double be = sc[S_BETA];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
        p[i] = z[i] + be * p[i];
}

__global__ void k_copy_skip(long long N, double* __restrict__ dst, const double* __restrict__ src,
                            const int* __restrict__ ctrl) {
    if (ctrl && ctrl[C_DONE]) return;
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
    if (ctrl && ctrl[C_DONE]) return;
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
    return g < 1 ? 1 : (g > kMaxBlocks ? kMaxBlocks : g);
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
    double* parts = w.partials.reserve(kMaxBlocks + 8);
    double* sc = w.scal.reserve(S_NUM + 8);
    int* ctrl = w.ctrl.reserve(C_NUM + 8);
    fill<int>(ctrl, 0, C_NUM, s);
    double host_sc[S_NUM] = {0};
    host_sc[S_TOL] = o.tol;
    SGIPC_HIP(hipMemcpyAsync(sc, host_sc, sizeof(host_sc), hipMemcpyHostToDevice, s));
    int max_iter = o.max_iter > 0 ? o.max_iter : (int)(o.max_iter_ratio * (double)N);
    const int gN = grid_for(N);
    fill<double>(x, 0.0, N, s);
    hipLaunchKernelGGL(k_copy_skip, dim3(gN), dim3(kThreads), 0, s, N, r, b, (const int*)nullptr);
    SGIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dot<OP_BNORM>, dim3(gN), dim3(kThreads), 0, s, N, b, b, parts, sc, ctrl);
    SGIPC_LAUNCH_CHECK();
    precond_apply_ctrl(c, r, z, nullptr);
    hipLaunchKernelGGL(k_dot<OP_INIT>, dim3(gN), dim3(kThreads), 0, s, N, r, z, parts, sc, ctrl);
    SGIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_copy_skip, dim3(gN), dim3(kThreads), 0, s, N, p, z, (const int*)nullptr);
    SGIPC_LAUNCH_CHECK();
    const Bsr& B = c.bsr;
    constexpr int G = 8;
    const int gS = std::max(1, std::min(kMaxBlocks, div_up((long long)B.nb * G, kThreads)));
    int check_every = o.check_every > 0 ? o.check_every : 4;
    int k = 1;
    int hctrl[C_NUM] = {0};
    bool done = false;
    for (; k < max_iter; ++k) {
        hipLaunchKernelGGL(k_spmv_pap<G>, dim3(gS), dim3(kThreads), 0, s, B.nb, B.rowptr.p, B.col.p, B.val.p, B.nnzb, p,
                           Ap, parts, sc, ctrl, k, o.stop_rule);
        SGIPC_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_update_xr, dim3(gN), dim3(kThreads), 0, s, N, x, r, p, Ap, parts, sc, ctrl, k, o.stop_rule);
        SGIPC_LAUNCH_CHECK();
        precond_apply_ctrl(c, r, z, ctrl);
        hipLaunchKernelGGL(k_dot<OP_BETA>, dim3(gN), dim3(kThreads), 0, s, N, r, z, parts, sc, ctrl);
        SGIPC_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_update_p, dim3(gN), dim3(kThreads), 0, s, N, p, z, sc, ctrl);
        SGIPC_LAUNCH_CHECK();
        if ((k % check_every) == 0 || k + 1 >= max_iter) {
            SGIPC_HIP(hipMemcpyAsync(hctrl, ctrl, sizeof(int) * C_NUM, hipMemcpyDeviceToHost, s));
            SGIPC_HIP(hipStreamSynchronize(s));
            if (hctrl[C_DONE]) { done = true; break; }
        }
    }
    SGIPC_HIP(hipMemcpyAsync(hctrl, ctrl, sizeof(int) * C_NUM, hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipMemcpyAsync(host_sc, sc, sizeof(host_sc), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    done = hctrl[C_DONE] != 0;
    st.converged = done ? 1 : 0;
    st.iterations = done ? hctrl[C_ITERS] : max_iter;
    st.rz0 = host_sc[S_RZ0];
    // true residual ||b - A x|| / ||b||, FP64 (tmp = b - A x)
    double* tmp = w.xw.reserve(N + 1);
    hipLaunchKernelGGL(k_copy_skip, dim3(gN), dim3(kThreads), 0, s, N, tmp, b, (const int*)nullptr);
    SGIPC_LAUNCH_CHECK();
    spmv64(c, -1.0, x, 1.0, tmp);
    fill<int>(ctrl, 0, C_NUM, s);
    hipLaunchKernelGGL(k_dot<OP_BNORM>, dim3(gN), dim3(kThreads), 0, s, N, tmp, tmp, parts, sc, ctrl);
    SGIPC_LAUNCH_CHECK();
    double rr = 0.0, bb = 0.0;
    SGIPC_HIP(hipMemcpyAsync(&rr, sc + S_BB, sizeof(double), hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(k_dot<OP_BNORM>, dim3(gN), dim3(kThreads), 0, s, N, b, b, parts, sc, ctrl);
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(hipMemcpyAsync(&bb, sc + S_BB, sizeof(double), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    st.rel_residual = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
}

}  // namespace sgipc
