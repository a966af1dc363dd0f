// Preconditioned conjugate gradients on the converted block system.
//
// Reference: S/linear_system/solver/pcg_solver.cu:158-267 (PCGSolver::pcg),
// dot products :8-150 (two-stage warp reduction + blocking cudaMemcpy per
// dot), vector updates :105-122; config pcg_solver.h:6-16.
//
// Here every scalar (rz, alpha, beta, the stop decision) lives in device
// memory: each producer of a dot product writes one partial per workgroup
// and every consumer workgroup sums the partials itself in a fixed order
// (deterministic, no atomics, no host round trip).  The iteration counter
// and the stop flag are device words too, so an iteration pair is a fixed
// launch sequence that is captured once into a HIP graph and replayed; the
// host only polls the stop flag every few iterations.  The iteration count
// and the iterate are those of the reference loop (stop test on the stale
// rz after the x/r update).
#include "spmv_sym.hpp"
#include "pcg_dev.hpp"
#include <atomic>
#include <chrono>
#include <thread>
#include <cmath>

namespace sgipc {
namespace {

constexpr int kThreads = 1024;    // workgroup size of the PCG kernels (16 waves)
constexpr int kRedBlocks = 1024;  // producers of dot partials use at most this many workgroups
constexpr int kCollectThreads = 256;  // MAS collect: one block row per thread, workgroups for every CU

// partial dot products of a.b, one per block (FP64 accumulation)
template <typename VA, typename VB>
__global__ void __launch_bounds__(kThreads) k_dot_part(long long N, const VA* __restrict__ a, const VB* __restrict__ b,
                                                       double* __restrict__ part, const int* __restrict__ ctrl) {
    if (ctrl && ctrl[C_STOP]) return;
    __shared__ double sh[kThreads / 64];
    double v = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
        v += (double)a[i] * (double)b[i];
    write_partial(v, part, sh);
}

// scalar = sum of partials into sc[idx] (one block)
__global__ void __launch_bounds__(kThreads) k_store_sum(const double* __restrict__ part, int np, double* __restrict__ sc,
                                                        int idx) {
    __shared__ double sh[kThreads / 64];
    double s = sum_partials(part, np, sh);
    if (threadIdx.x == 0) sc[idx] = s;
}

// Direction update fused with the SpMV.  With beta = rz_(k-1) / rz_(k-2)
// (0 at k = 1): p_k = z + beta p_(k-1) and, by linearity,
// A p_k = A z + beta A p_(k-1), so the kernel gathers only z (symmetric
// SpMV over the upper BCOO + transposed references), forms p and Ap row by
// row in place, and writes per-workgroup partials of p.Ap.  First kernel of
// an iteration: enforces the iteration budget (iteration k runs while
// k < max_iter).
template <int G, typename MV, typename V>
__global__ void __launch_bounds__(kThreads) k_spmv_pap(int nb, const int* __restrict__ eptr, const int2* __restrict__ ent,
                                                       const MV* __restrict__ val, const V* __restrict__ z,
                                                       V* __restrict__ p, V* __restrict__ Ap,
                                                       const double* __restrict__ rz_last,
                                                       const double* __restrict__ rz_prev, int np_rz,
                                                       double* __restrict__ part, int* __restrict__ ctrl) {
    if (ctrl[C_STOP]) return;
    const int iter = ctrl[C_ITER];
    if (iter >= ctrl[C_MAXIT]) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { ctrl[C_NOCONV] = 1; ctrl[C_STOP] = ctrl[C_MAXIT]; }
        return;
    }
    __shared__ double sh[2 * (kThreads / 64)];
    const bool first = iter == 1;
    double be = 0.0;
    if (!first) {
        double rn, ro;
        sum_partials2(rz_last, np_rz, rz_prev, np_rz, sh, rn, ro);
        be = rn / ro;
    }
    auto zv = [z](size_t i) { return z[i]; };
    constexpr int RPB = kThreads / G;
    const int nchunks = (nb + RPB - 1) / RPB;
    ChunkWalk w = xcd_chunks(nchunks);
    double contrib = 0.0;
    for (int ch = w.first; ch < w.end; ch += w.step) {
        int row = ch * RPB + threadIdx.x / G, sub = threadIdx.x % G;
        MV s0 = 0, s1 = 0, s2 = 0;
        if (row < nb) sym_row_lane<G, MV>(row, sub, eptr, ent, val, zv, s0, s1, s2);
        group_sum3<G, MV>(s0, s1, s2);  // every lane of the group holds the row sums
        if (sub < 3 && row < nb) {        // lane q forms component q
            const size_t o = 3 * (size_t)row + sub;
            const V aw = (V)(sub == 0 ? s0 : (sub == 1 ? s1 : s2));
            V pq, aq;
            if (first) {
                pq = z[o];
                aq = aw;
            } else {
                pq = (V)((double)z[o] + be * (double)p[o]);
                aq = (V)((double)aw + be * (double)Ap[o]);
            }
            p[o] = pq;
            Ap[o] = aq;
            contrib += (double)pq * (double)aq;
        }
    }
    write_partial(contrib, part, sh);
}

// alpha = rz / p.Ap ; x += alpha p ; r -= alpha Ap ; partials of r.r ; then
// the iteration counter advances (read again only by the next iteration's
// kernels and the stop tests, which report ctrl[C_ITER] - 1).
template <typename V>
__global__ void __launch_bounds__(kThreads) k_update_xr(long long N, double* __restrict__ x, V* __restrict__ r,
                                                        const V* __restrict__ p, const V* __restrict__ Ap,
                                                        const double* __restrict__ part_pap, int np_pap,
                                                        const double* __restrict__ part_rz, int np_rz,
                                                        double* __restrict__ part_rr, int* __restrict__ ctrl) {
    if (ctrl[C_STOP]) return;
    __shared__ double sh[2 * (kThreads / 64)];
    double rz, pap;
    sum_partials2(part_rz, np_rz, part_pap, np_pap, sh, rz, pap);
    double al = rz / pap;
    double v = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        x[i] += al * (double)p[i];
        V ri = (V)((double)r[i] - al * (double)Ap[i]);
        r[i] = ri;
        v += (double)ri * (double)ri;
    }
    write_partial(v, part_rr, sh);
    if (blockIdx.x == 0 && threadIdx.x == 0) ctrl[C_ITER] += 1;
}

// Stop test, then z = sum_l Z_l[node_l(v)] on the MAS range, the 12x12
// block-Jacobi of the affine bodies on block rows < abd_rows (when abd_inv is
// set), identity elsewhere, and partials of r.z; one thread per block row.
template <typename V, typename T>
__global__ void __launch_bounds__(kCollectThreads) k_collect_dot(int nb, const V* __restrict__ r, V* __restrict__ z,
                                                          int sub_off, int n_mas, int L, const int* __restrict__ node_of,
                                                          const T* __restrict__ Z, const double* __restrict__ abd_inv,
                                                          int abd_rows, double* __restrict__ part_rz,
                                                          const double* __restrict__ part_rz_prev, int np_rz,
                                                          const double* __restrict__ part_rr, int np_rr,
                                                          double* __restrict__ sc, int* __restrict__ ctrl, int rule,
                                                          volatile int* hp, volatile double* hq) {
    if (ctrl[C_STOP]) {
        if (hp && blockIdx.x == 0 && threadIdx.x == 0) publish_progress(ctrl, sc, hp, hq);
        return;
    }
    __shared__ double sh[kCollectThreads / 64];
    double v = 0.0;
    // the stop test's partials are loaded first and reduced after z (whose
    // loads do not depend on the test), so the two latencies overlap
    const double share = stop_share(rule, part_rz_prev, np_rz, part_rr, np_rr);
    for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < nb; row += gridDim.x * blockDim.x) {
        size_t o = 3 * (size_t)row;
        V r0 = r[o], r1 = r[o + 1], r2 = r[o + 2];
        V z0, z1, z2;
        int vt = row - sub_off;
        if (vt >= 0 && vt < n_mas) {
            // levels eight at a time: all node ids, then all Z values, are
            // loaded before the sums (kept in level order)
            T s0 = 0, s1 = 0, s2 = 0;
            for (int l0 = 0; l0 < L; l0 += kLevelBatch) {
                size_t g[kLevelBatch];
#pragma unroll
                for (int k = 0; k < kLevelBatch; ++k)
                    g[k] = (l0 + k < L) ? 3 * (size_t)node_of[(size_t)(l0 + k) * n_mas + vt] : 0;
                T za[kLevelBatch], zb[kLevelBatch], zc[kLevelBatch];
#pragma unroll
                for (int k = 0; k < kLevelBatch; ++k)
                    if (l0 + k < L) { za[k] = Z[g[k]]; zb[k] = Z[g[k] + 1]; zc[k] = Z[g[k] + 2]; }
#pragma unroll
                for (int k = 0; k < kLevelBatch; ++k)
                    if (l0 + k < L) { s0 += za[k]; s1 += zb[k]; s2 += zc[k]; }
            }
            z0 = (V)s0; z1 = (V)s1; z2 = (V)s2;
        } else if (row < abd_rows) {
            const int b = row >> 2, k = row & 3;
            const double* P = abd_inv + 144 * (size_t)b + 36 * k;  // rows 3k .. 3k+2 of the body inverse
            const V* rb = r + 12 * (size_t)b;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll 2  // a full unroll's 36 loads in flight would set the whole kernel's register budget
            for (int j = 0; j < 12; ++j) {
                const double rj = (double)rb[j];
                s0 += P[j] * rj; s1 += P[12 + j] * rj; s2 += P[24 + j] * rj;
            }
            z0 = (V)s0; z1 = (V)s1; z2 = (V)s2;
        } else {
            z0 = r0; z1 = r1; z2 = r2;
        }
        z[o] = z0; z[o + 1] = z1; z[o + 2] = z2;
        v += (double)r0 * (double)z0 + (double)r1 * (double)z1 + (double)r2 * (double)z2;
    }
    const bool stop = stop_test_share(rule, share, sc, ctrl, sh);
    // the thread that ran the test's writes publishes its outcome
    if (hp && blockIdx.x == 0 && threadIdx.x == 0) publish_progress(ctrl, sc, hp, hq);
    if (stop) return;
    write_partial(v, part_rz, sh);
}

// Stop test, then z = D^-1 r (block Jacobi) or z = r, and partials of r.z
template <typename V>
__global__ void __launch_bounds__(kThreads) k_simple_precond_dot(long long N, const V* __restrict__ r, V* __restrict__ z,
                                                                 const double* __restrict__ Dinv, double* __restrict__ part_rz,
                                                                 const double* __restrict__ part_rz_prev, int np_rz,
                                                                 const double* __restrict__ part_rr, int np_rr,
                                                                 double* __restrict__ sc, int* __restrict__ ctrl,
                                                                 int rule) {
    if (ctrl[C_STOP]) return;
    __shared__ double sh[kThreads / 64];
    double v = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        V zi;
        if (Dinv) {
            long long b = i / 3;
            int p = (int)(i % 3);
            const double* D = Dinv + 9 * (size_t)b;
            zi = (V)(D[p] * (double)r[3 * b] + D[p + 3] * (double)r[3 * b + 1] + D[p + 6] * (double)r[3 * b + 2]);
        } else {
            zi = r[i];
        }
        z[i] = zi;
        v += (double)r[i] * (double)zi;
    }
    if (stop_test(rule, part_rz_prev, np_rz, part_rr, np_rr, sc, ctrl, sh)) return;
    write_partial(v, part_rz, sh);
}

// The same with the row-contiguous FP32 copy of both triangles: the G lanes
// of a row read consecutive entries (40 bytes each: two 16-byte loads at
// 8-byte alignment and one 8-byte load that also brings the column), no
// block index, no transposition in registers, no column stream (44 -> 40
// bytes per entry: counted traffic 96.7 -> 80.4 MB per launch at 137 k
// block rows, the launch time within box-to-box noise of 2-3 %); the
// column's three z values come in one 12-byte load; same
// entry order and the same products, so the same sums.  (Measured and kept
// out: 9-float entries read at 4-byte alignment, 30 % slower; the row's own
// z, p, Ap loaded ahead of the row sums, 35 % slower.)
template <int G, typename V>
__global__ void __launch_bounds__(kThreads) k_spmv_pap_full(int nb, const int* __restrict__ eptr,
                                                            const float* __restrict__ vf, const V* __restrict__ z,
                                                            V* __restrict__ p, V* __restrict__ Ap,
                                                            const double* __restrict__ rz_last,
                                                            const double* __restrict__ rz_prev, int np_rz,
                                                            double* __restrict__ part, int* __restrict__ ctrl) {
    if (ctrl[C_STOP]) return;
    const int iter = ctrl[C_ITER];
    if (iter >= ctrl[C_MAXIT]) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { ctrl[C_NOCONV] = 1; ctrl[C_STOP] = ctrl[C_MAXIT]; }
        return;
    }
    __shared__ double sh[2 * (kThreads / 64)];
    const bool first = iter == 1;
    // beta = rz_last / rz_prev is formed after the first chunk's row sums,
    // so its partial-sum reduction overlaps the matrix loads (the chunk loop
    // is uniform over the workgroup, so every thread reaches the barrier).
    // The FP64 kernel keeps it first: holding its row state across the
    // reduction cost it occupancy.
    double be = 0.0;
    bool have_be = first;
    constexpr int RPB = kThreads / G;
    const int nchunks = (nb + RPB - 1) / RPB;
    ChunkWalk w = xcd_chunks(nchunks);
    double contrib = 0.0;
    for (int ch = w.first; ch < w.end; ch += w.step) {
        int row = ch * RPB + threadIdx.x / G, sub = threadIdx.x % G;
        float s0 = 0, s1 = 0, s2 = 0;
        if (row < nb) {
            const int end = eptr[row + 1];
            int e = eptr[row] + sub;
            for (; e + G < end; e += 2 * G) {  // two entries in flight
                const FullEntry fa = load_full_entry(vf, e), fb = load_full_entry(vf, e + G);
                const f3_a4 xa = load_z3(z, fa.col), xb = load_z3(z, fb.col);
                full_madd(fa, xa, s0, s1, s2);
                full_madd(fb, xb, s0, s1, s2);
            }
            if (e < end) {
                const FullEntry f = load_full_entry(vf, e);
                full_madd(f, load_z3(z, f.col), s0, s1, s2);
            }
        }
        group_sum3<G, float>(s0, s1, s2);
        if (!have_be) {
            double rn, ro;
            sum_partials2(rz_last, np_rz, rz_prev, np_rz, sh, rn, ro);
            be = rn / ro;
            have_be = true;
        }
        if (sub < 3 && row < nb) {
            const size_t o = 3 * (size_t)row + sub;
            const V aw = (V)(sub == 0 ? s0 : (sub == 1 ? s1 : s2));
            V pq, aq;
            if (first) {
                pq = z[o];
                aq = aw;
            } else {
                pq = (V)((double)z[o] + be * (double)p[o]);
                aq = (V)((double)aw + be * (double)Ap[o]);
            }
            p[o] = pq;
            Ap[o] = aq;
            contrib += (double)pq * (double)aq;
        }
    }
    write_partial(contrib, part, sh);
}

template <typename TD, typename TS>
__global__ void k_copy_cast(long long N, TD* __restrict__ dst, const TS* __restrict__ src) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
        dst[i] = (TD)src[i];
}

__global__ void k_axpy(long long N, double* __restrict__ y, double a, const double* __restrict__ x) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}


__global__ void k_ctrl_init(int* ctrl, int maxit) {
    ctrl[C_STOP] = 0;
    ctrl[C_ITER] = 1;
    ctrl[C_MAXIT] = maxit;
    ctrl[C_NOCONV] = 0;
    ctrl[C_PUB] = 0;
}

// Last kernel of a captured group of iterations: copies the stop and
// iteration words and the progress ratio to pinned host memory, then (after
// a system-scope fence) the publication count the host waits on.
__global__ void k_publish(int* __restrict__ ctrl, const double* __restrict__ sc, volatile int* hp,
                          volatile double* hq) {
    publish_progress(ctrl, sc, hp, hq);
}

__global__ void k_set_scalar(double* sc, int idx, double v) { sc[idx] = v; }

// rz0 = sum of the r.z partials; a zero rz0 (zero right-hand side) stops the
// loop before its first iteration, so x stays 0 (the reference would divide
// 0/0), with no host round trip: the caller reads rz0 with the final words.
__global__ void __launch_bounds__(kThreads) k_store_rz0(const double* __restrict__ part, int np,
                                                        double* __restrict__ sc, int* __restrict__ ctrl) {
    __shared__ double sh[kThreads / 64];
    const double s = sum_partials(part, np, sh);
    if (threadIdx.x == 0) {
        sc[S_RZ0] = s;
        if (s == 0.0) ctrl[C_STOP] = 1;
    }
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

// the affine-body block-Jacobi applies on the body rows below the MAS range
const double* abd_inv(const Ctx& c) {
    return (c.abd.inv_ready && c.sub_offset >= 4 * c.abd.nb) ? (const double*)c.abd.inv.p : nullptr;
}
int abd_rows(const Ctx& c) { return abd_inv(c) ? 4 * c.abd.nb : 0; }

int grid_for(long long N) {
    int g = div_up(N, kThreads);
    return g < 1 ? 1 : (g > kRedBlocks ? kRedBlocks : g);
}

// ---- the PCG engine --------------------------------------------------------
// Vectors of type V (double or float), matrix values MV, MAS data T.
template <typename V, typename MV, typename T>
struct Engine {
    Ctx& c;
    long long N;
    int gN, gS;
    V *r, *z, *p, *Ap;
    double* x;
    double *part_rz[2], *part_pap, *part_rr, *part_bb;
    double* sc;
    int* ctrl;
    const MV* val;
    bool mas;
    T *R, *Z;
    const double* Dinv;
    int rule;
    int maxit = 0;
    double tol_now = 0.0;  // tolerance of the current run (progress prediction)
    int gU = 1;  // workgroups (and r.r partials) of the x/r update
    int nRZ = 1; // r.z partials written by the preconditioner apply
    bool fused = false;  // MAS apply as one launch

    int G = 8;  // lanes per block row in the SpMV (4, 8 or 16)
    double rz0_h = 0.0;  // rz0 of the last run, read with its final words
    // mixed precision: the row-contiguous FP32 copy (null: the upper BCOO)
    const float* vfull = nullptr;
    // pinned progress words published by the MAS collect kernel (set while
    // the last iteration of a captured group is enqueued)
    volatile int* pub_h = nullptr;
    volatile double* pub_q = nullptr;
    static constexpr bool kFp32Mas = sizeof(T) == 4;

    // z = M^-1 r fused with the r.z partials (nRZ workgroups) and the stop test
    void precond_dot(double* rz_out, const double* rz_prev, int rule_now) {
        hipStream_t s = c.stream;
        if (mas) {
            if (fused) mas_solve_with_top(c, kFp32Mas, ctrl);
            else mas_bank_solve(c, kFp32Mas, ctrl);
            hipLaunchKernelGGL((k_collect_dot<V, T>), dim3(nRZ), dim3(kCollectThreads), 0, s, c.nb, r, z, c.sub_offset,
                               c.mas.n, c.mas.L, c.mas.node_of.p, (const T*)Z, abd_inv(c), abd_rows(c), rz_out,
                               rz_prev, nRZ, part_rr, gU, sc, ctrl, rule_now, pub_h, pub_q);
            SGIPC_LAUNCH_CHECK();
        } else {
            hipLaunchKernelGGL((k_simple_precond_dot<V>), dim3(gN), dim3(kThreads), 0, s, N, r, z, Dinv, rz_out,
                               rz_prev, nRZ, part_rr, gU, sc, ctrl, rule_now);
            SGIPC_LAUNCH_CHECK();
        }
    }

    // kernels of iteration k with parity par = k & 1
    void enqueue_iteration(int par) {
        hipStream_t s = c.stream;
        const SymBcoo& S = c.bsr;
        double* rz_last = part_rz[par ^ 1];  // rz_(k-1)
        double* rz_this = part_rz[par];      // holds rz_(k-2); receives rz_k
        prof_begin(c, PROF_SPMV);
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(gS), dim3(kThreads), 0, s, S.nb, S.eptr.p, S.ent.p, val, (const V*)z, p, Ap,
                               (const double*)rz_last, (const double*)rz_this, nRZ, part_pap, ctrl);
        };
        if (vfull) {
            auto launchf = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(gS), dim3(kThreads), 0, s, S.nb, S.eptr.p, vfull, (const V*)z, p,
                                   Ap, (const double*)rz_last, (const double*)rz_this, nRZ, part_pap, ctrl);
            };
            if (G == 4) launchf(k_spmv_pap_full<4, V>);
            else if (G == 16) launchf(k_spmv_pap_full<16, V>);
            else launchf(k_spmv_pap_full<8, V>);
        } else if (G == 4) launch(k_spmv_pap<4, MV, V>);
        else if (G == 16) launch(k_spmv_pap<16, MV, V>);
        else launch(k_spmv_pap<8, MV, V>);
        SGIPC_LAUNCH_CHECK();
        prof_end(c, PROF_SPMV);
        if (mas) {
            mas_update_restrict(c, x, r, p, Ap, part_pap, gS, rz_last, nRZ, part_rr, ctrl, kFp32Mas, !fused);
        } else {
            hipLaunchKernelGGL((k_update_xr<V>), dim3(gN), dim3(kThreads), 0, s, N, x, r, (const V*)p, (const V*)Ap,
                               (const double*)part_pap, gS, (const double*)rz_last, nRZ, part_rr, ctrl);
            SGIPC_LAUNCH_CHECK();
        }
        precond_dot(rz_this, rz_last, rule);
    }

    // r holds the right-hand side, x the initial guess is zero.  rz0 stays on
    // the device (run() reads it with its final words into rz0_h; 0 means a
    // zero right-hand side, for which the loop stops at once).
    void init(double tol, int max_iter) {
        hipStream_t s = c.stream;
        maxit = max_iter;
        tol_now = tol;
        if (mas) {
            int g1, gx;
            mas_update_layout(c, g1, gx);
            gU = g1 + gx;
        } else {
            gU = gN;
        }
        fused = mas && mas_fused(c);
        // r.z partials: one per workgroup of the collect (256-thread
        // workgroups, so a system of ~100 k block rows spreads over every CU)
        // or of the simple preconditioner apply
        nRZ = mas ? std::max(1, std::min(div_up(c.nb, kCollectThreads), kRedBlocks)) : gN;
        hipLaunchKernelGGL(k_ctrl_init, dim3(1), dim3(1), 0, s, ctrl, max_iter);
        SGIPC_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, s, sc, (int)S_TOL, tol);
        SGIPC_LAUNCH_CHECK();
        fill<double>(x, 0.0, N, s);
        hipLaunchKernelGGL((k_dot_part<V, V>), dim3(gN), dim3(kThreads), 0, s, N, r, r, part_bb, (const int*)nullptr);
        SGIPC_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_store_sum, dim3(1), dim3(kThreads), 0, s, part_bb, gN, sc, (int)S_BB);
        SGIPC_LAUNCH_CHECK();
        if (mas) mas_restrict(c, r, 3LL * c.sub_offset, kFp32Mas, ctrl, !fused);
        precond_dot(part_rz[0], part_rz[1], -1);
        hipLaunchKernelGGL(k_store_rz0, dim3(1), dim3(kThreads), 0, s, part_rz[0], nRZ, sc, ctrl);
        SGIPC_LAUNCH_CHECK();
    }

    size_t graph_key() const {
        size_t h = 1469598103934665603ull;
        auto mix = [&](size_t v) { h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2); };
        const SymBcoo& S = c.bsr;
        mix((size_t)N); mix((size_t)gN); mix((size_t)gS); mix((size_t)G); mix((size_t)gU); mix((size_t)S.nb); mix((size_t)S.K);
        mix((size_t)S.eptr.p); mix((size_t)S.ent.p); mix((size_t)val); mix((size_t)r); mix((size_t)z); mix((size_t)p);
        mix((size_t)Ap); mix((size_t)x); mix((size_t)part_rz[0]); mix((size_t)sc);
        mix((size_t)ctrl); mix((size_t)mas); mix((size_t)R); mix((size_t)Z); mix((size_t)Dinv); mix((size_t)rule);
        mix((size_t)c.sub_offset); mix((size_t)c.nb); mix((size_t)c.stream);
        mix((size_t)abd_inv(c)); mix((size_t)abd_rows(c)); mix((size_t)vfull);
        mix((size_t)fused); mix((size_t)nRZ);
        if (mas) {
            const MasState& m = c.mas;
            mix((size_t)m.nbanks); mix((size_t)m.L); mix((size_t)m.M); mix((size_t)m.n); mix((size_t)m.bank);
            mix((size_t)m.nb1); mix((size_t)m.going_next.p); mix((size_t)m.node_of.p); mix((size_t)m.level_size.p);
            mix((size_t)m.part_map.p); mix((size_t)m.psum0.p); mix((size_t)m.l1lo.p); mix((size_t)m.l1hi.p);
            mix((size_t)m.comp_mask.p); mix((size_t)m.fine_mask.p);
            mix((size_t)m.inv64.p); mix((size_t)m.inv32.p); mix((size_t)m.R64.p); mix((size_t)m.R32.p);
            mix((size_t)m.Z64.p); mix((size_t)m.Z32.p);
            mix((size_t)m.inv16_ok); mix((size_t)m.inv16.p); mix((size_t)m.inv16_d.p);
        }
        return h;
    }

    static constexpr int kBigGroup = 8;  // iterations in the long captured group (even: parities repeat)

    // Capture n (even) iterations into a graph; with pinned polling the
    // group's last collect kernel (or k_publish) publishes the progress.
    hipGraphExec_t capture_group(int n, bool pinned, PinnedPoll& pp) {
        hipStream_t s = c.stream;
        hipGraph_t g = nullptr;
        SGIPC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            for (int i = 0; i < n; ++i) {
                if (i == n - 1 && pinned && mas) {  // the group's last collect kernel publishes
                    pub_h = pp.d;
                    pub_q = reinterpret_cast<volatile double*>(pp.d + 8);
                }
                enqueue_iteration((i + 1) & 1);
            }
            pub_h = nullptr;
            pub_q = nullptr;
            if (pinned && !mas) {
                hipLaunchKernelGGL(k_publish, dim3(1), dim3(1), 0, s, ctrl, (const double*)sc, pp.d,
                                   reinterpret_cast<double*>(pp.d + 8));
                SGIPC_LAUNCH_CHECK();
            }
        } catch (...) {
            pub_h = nullptr;
            pub_q = nullptr;
            hipGraph_t dead = nullptr;
            (void)hipStreamEndCapture(s, &dead);
            if (dead) (void)hipGraphDestroy(dead);
            throw;
        }
        SGIPC_HIP(hipStreamEndCapture(s, &g));
        hipGraphExec_t exec = nullptr;
        hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        SGIPC_HIP(e);
        return exec;
    }

    // Run iterations until the device stop flag is set; returns the
    // reference iteration count and whether it converged.
    int run(PcgGraph& gcache, int check_every, bool& converged) {
        hipStream_t s = c.stream;
        if (maxit <= 1) {  // no iteration fits the budget
            SGIPC_HIP(hipMemcpyAsync(&rz0_h, sc + S_RZ0, sizeof(double), hipMemcpyDeviceToHost, s));
            SGIPC_HIP(hipStreamSynchronize(s));
            converged = false;
            return maxit;
        }
        // profiling records events per kernel, which a captured graph cannot
        const bool use_graph = !c.prof.on;
        int pairs_per_check = std::max(1, check_every / 2);
        PinnedPoll& pp = c.pcg.poll;
        if (use_graph && !pp.h && !pp.failed) {
            void* hptr = nullptr;
            if (hipHostMalloc(&hptr, 64, hipHostMallocMapped | hipHostMallocCoherent) == hipSuccess) {
                void* dptr = nullptr;
                if (hipHostGetDevicePointer(&dptr, hptr, 0) == hipSuccess) {
                    pp.h = static_cast<int*>(hptr);
                    pp.d = static_cast<int*>(dptr);
                } else {
                    (void)hipHostFree(hptr);
                    pp.failed = true;
                }
            } else {
                pp.failed = true;
            }
            (void)hipGetLastError();
        }
        const bool pinned = use_graph && pp.h;
        if (use_graph) {
            size_t key = graph_key();
            if (!gcache.exec || gcache.key != key) {
                if (gcache.exec) { (void)hipGraphExecDestroy(gcache.exec); gcache.exec = nullptr; }
                if (gcache.exec_big) { (void)hipGraphExecDestroy(gcache.exec_big); gcache.exec_big = nullptr; }
                gcache.exec = capture_group(2, pinned, pp);
                if (pinned) gcache.exec_big = capture_group(kBigGroup, pinned, pp);
                gcache.key = key;
            }
        }
        int hc[C_NUM] = {0};
        if (pinned) {
            // Keep two groups of iterations queued ahead of the device and
            // follow its published progress in pinned memory: the device never
            // waits for the host, and at most two groups run past the stop
            // (their kernels return at once).  Groups are kBigGroup
            // iterations long (one graph launch and one publication per
            // group) while the convergence rate seen in the publications
            // predicts more than that many iterations to go, two iterations
            // near the stop, so few iterations run past it.
            volatile int* hp = pp.h;
            volatile double* hq = reinterpret_cast<volatile double*>(pp.h + 8);
            hp[0] = 0; hp[1] = 0; hp[2] = 0;
            std::atomic_thread_fence(std::memory_order_seq_cst);
            constexpr int kAhead = 2;
            int launched = 0, last_done = -1;
            long long launched_iters = 0;
            int k_prev = -1;
            double q_prev = 0.0, stop_pred = -1.0;  // predicted stopping iteration (< 0: none yet)
            auto last_change = std::chrono::steady_clock::now();
            for (;;) {
                const int done = hp[2];
                std::atomic_thread_fence(std::memory_order_acquire);
                if (done > 0 && hp[0] != 0) break;  // stopped (converged or budget)
                if (done != last_done && done > 0) {  // a new publication: update the prediction
                    const int k = hp[1] - 1;
                    const double q = hq[0];
                    if (k_prev >= 0 && k > k_prev && q > 0.0 && q_prev > q && tol_now > 0.0) {
                        const double rate = std::log(q / q_prev) / (double)(k - k_prev);  // < 0
                        stop_pred = (double)k + std::log(tol_now / q) / rate;
                    }
                    k_prev = k;
                    q_prev = q;
                }
                if (launched_iters < (long long)maxit + 2 && launched - done < kAhead) {
                    const bool big = gcache.exec_big &&
                                     (stop_pred < 0.0 || stop_pred - (double)launched_iters > kBigGroup + 2);
                    SGIPC_HIP(hipGraphLaunch(big ? gcache.exec_big : gcache.exec, s));
                    launched_iters += big ? kBigGroup : 2;
                    ++launched;
                    if (done != last_done) {
                        last_done = done;
                        last_change = std::chrono::steady_clock::now();
                    }
                    continue;
                }
                if (done >= launched) break;        // all launched groups ran, none stopped: budget used up
                if (done != last_done) {
                    last_done = done;
                    last_change = std::chrono::steady_clock::now();
                } else if (std::chrono::steady_clock::now() - last_change > std::chrono::seconds(2)) {
                    // no progress for 2 s: let the stream report its state
                    const hipError_t q = hipStreamQuery(s);
                    if (q != hipSuccess && q != hipErrorNotReady) SGIPC_HIP(q);
                    if (q == hipSuccess) break;     // idle: read the words below
                    last_change = std::chrono::steady_clock::now();
                }
#if defined(__x86_64__)
                __builtin_ia32_pause();
#else
                std::this_thread::yield();
#endif
            }
            SGIPC_HIP(hipMemcpyAsync(hc, ctrl, sizeof(int) * C_NUM, hipMemcpyDeviceToHost, s));
            SGIPC_HIP(hipMemcpyAsync(&rz0_h, sc + S_RZ0, sizeof(double), hipMemcpyDeviceToHost, s));
            SGIPC_HIP(hipStreamSynchronize(s));
        } else if (use_graph) {
            // Poll the stop word; after two polls the next one is placed where
            // the measured convergence rate predicts the stop (at most twice
            // the nominal interval ahead), so few polls are needed and few
            // iterations run past the stop.
            int npairs = pairs_per_check, k_prev = 0;
            double q_prev = 0.0, q = 0.0;
            for (;;) {
                for (int j = 0; j < npairs; ++j) SGIPC_HIP(hipGraphLaunch(gcache.exec, s));
                SGIPC_HIP(hipMemcpyAsync(hc, ctrl, sizeof(int) * C_NUM, hipMemcpyDeviceToHost, s));
                SGIPC_HIP(hipMemcpyAsync(&q, sc + S_PROG, sizeof(double), hipMemcpyDeviceToHost, s));
                SGIPC_HIP(hipMemcpyAsync(&rz0_h, sc + S_RZ0, sizeof(double), hipMemcpyDeviceToHost, s));
                SGIPC_HIP(hipStreamSynchronize(s));
                if (hc[C_STOP]) break;
                const int k = hc[C_ITER] - 1;  // iterations done; q is the ratio after iteration k
                npairs = pairs_per_check;
                if (k_prev > 0 && k > k_prev && q > 0.0 && q_prev > q && tol_now > 0.0) {
                    const double rate = std::log(q / q_prev) / (double)(k - k_prev);  // < 0
                    const double need = std::log(tol_now / q) / rate;                  // iterations left
                    npairs = std::max(1, std::min(2 * pairs_per_check, (int)std::ceil(need / 2.0)));
                }
                k_prev = k;
                q_prev = q;
            }
        } else {
            // timed (eager) mode: same batching; the event records of launches
            // after the stop (no work done) are dropped afterwards
            const size_t base_spmv = c.prof.pend[PROF_SPMV].size();
            const size_t base_mas = c.prof.pend[PROF_MAS_SOLVE].size();
            for (;;) {
                for (int q = 0; q < pairs_per_check; ++q) {
                    enqueue_iteration(1);
                    enqueue_iteration(0);
                }
                SGIPC_HIP(hipMemcpyAsync(hc, ctrl, sizeof(int) * C_NUM, hipMemcpyDeviceToHost, s));
                SGIPC_HIP(hipMemcpyAsync(&rz0_h, sc + S_RZ0, sizeof(double), hipMemcpyDeviceToHost, s));
                SGIPC_HIP(hipStreamSynchronize(s));
                if (hc[C_STOP]) break;
            }
            const size_t useful = (size_t)(hc[C_NOCONV] ? hc[C_STOP] - 1 : hc[C_STOP]);
            prof_drop_tail(c, PROF_SPMV, base_spmv + useful);
            if (mas) prof_drop_tail(c, PROF_MAS_SOLVE, base_mas + useful);
        }
        converged = hc[C_NOCONV] == 0;
        return hc[C_STOP];
    }
};

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

// z = M^-1 r with the configured preconditioner (standalone, outside PCG)
void precond_apply(Ctx& c, const double* r, double* z) {
    long long N = 3LL * c.nb;
    hipStream_t s = c.stream;
    if (c.precond_kind == 2 || c.precond_kind == 3) {
        long long off = 3LL * c.sub_offset, len = 3LL * c.mas.n;
        if (off > 0 || off + len < N) {
            hipLaunchKernelGGL((k_copy_cast<double, double>), dim3(grid_for(N)), dim3(kThreads), 0, s, N, z, r);
            SGIPC_LAUNCH_CHECK();
        }
        if (c.precond_kind == 2) mas_apply64(c, r + off, z + off);
        else mas_apply32(c, r + off, z + off);
        if (abd_inv(c)) abd_precond_apply(c, r, z);
        return;
    }
    double* part = c.pcg.partials.reserve(5 * kPart);
    int* ctrl = c.pcg.ctrl.reserve(C_NUM + 8);
    fill<int>(ctrl, 0, C_NUM, s);
    hipLaunchKernelGGL((k_simple_precond_dot<double>), dim3(grid_for(N)), dim3(kThreads), 0, s, N, r, z,
                       c.precond_kind == 1 ? (const double*)c.pcg.Dinv.p : (const double*)nullptr, part, part, 1, part,
                       1, (double*)nullptr, ctrl, -1);
    SGIPC_LAUNCH_CHECK();
}

// ||v||^2 into the device scalar sc[idx] (no host read)
static void device_norm2_to(Ctx& c, const double* v, long long N, double* part, double* sc, int idx) {
    hipStream_t s = c.stream;
    int g = grid_for(N);
    hipLaunchKernelGGL((k_dot_part<double, double>), dim3(g), dim3(kThreads), 0, s, N, v, v, part, (const int*)nullptr);
    SGIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_store_sum, dim3(1), dim3(kThreads), 0, s, part, g, sc, idx);
    SGIPC_LAUNCH_CHECK();
}
// device scalars sc[i0] and sc[i1] to the host in one synchronisation
static void read_scalars(Ctx& c, const double* sc, int i0, int i1, double& a, double& b) {
    const int lo = std::min(i0, i1), hi = std::max(i0, i1);
    double h[16];
    SGIPC_HIP(hipMemcpyAsync(h, sc + lo, sizeof(double) * (hi - lo + 1), hipMemcpyDeviceToHost, c.stream));
    SGIPC_HIP(hipStreamSynchronize(c.stream));
    a = h[i0 - lo];
    b = h[i1 - lo];
}

void pcg_solve(Ctx& c, const double* b, double* x, const SolveOpts& o, SolveStats& st) {
    if (!c.have_bcoo) throw state_error("solve: convert first");
    if (!c.have_bsr) bsr_build(c);
    hipStream_t s = c.stream;
    const long long N = 3LL * c.nb;
    PcgWork& w = c.pcg;
    const bool mas = (c.precond_kind == 2 || c.precond_kind == 3);
    if (mas && !c.mas.ready) throw state_error("solve: MAS preconditioner not built");
    if (c.precond_kind == 1 && !w.Dinv.p) throw state_error("solve: block-Jacobi preconditioner not built");
    double* parts = w.partials.reserve(6 * kPart);
    double* sc = w.scal.reserve(S_NUM + 8);
    int* ctrl = w.ctrl.reserve(C_NUM + 8);
    const int max_iter = o.max_iter > 0 ? o.max_iter : (int)(o.max_iter_ratio * (double)N);
    const int gN = grid_for(N);
    // SpMV: chunks of kThreads / G block rows, a multiple of 8 workgroups (one range per XCD)
    const int G = c.spmv_group;
    const int nchunks = (int)div_up((long long)c.nb, kThreads / G);
    int gS = std::min(kPart, std::max(8, (int)div_up(nchunks, 8) * 8));
    // the FP32 full-copy SpMV of mixed solves runs best as one resident round
    // of workgroups walking the chunks (the FP64 kernel as one per chunk)
    if (o.precision != 0 && c.spmv_blocks > 0) gS = std::min(gS, std::max(8, c.spmv_blocks / 8 * 8));
    st.refinements = 0;
    if (o.precision == 0) {
        auto run64 = [&](auto tag) {
            using T = decltype(tag);
            const bool fp32mas = sizeof(T) == 4;
            Engine<double, double, T> E{c, N, gN, gS,
                                        w.r.reserve(N + 1), w.z.reserve(N + 1), w.p.reserve(N + 1),
                                        w.Ap.reserve(N + 1),
                                        x, {parts, parts + kPart}, parts + 2 * kPart, parts + 3 * kPart,
                                        parts + 4 * kPart, sc, ctrl, (const double*)c.t_vals, mas,
                                        mas ? (T*)(fp32mas ? (void*)c.mas.R32.p : (void*)c.mas.R64.p) : nullptr,
                                        mas ? (T*)(fp32mas ? (void*)c.mas.Z32.p : (void*)c.mas.Z64.p) : nullptr,
                                        c.precond_kind == 1 ? w.Dinv.p : nullptr, o.stop_rule};
            E.G = G;
            hipLaunchKernelGGL((k_copy_cast<double, double>), dim3(gN), dim3(kThreads), 0, s, N, E.r, b);
            SGIPC_LAUNCH_CHECK();
            E.init(o.tol, max_iter);
            bool conv = false;
            const int it = E.run(fp32mas ? w.graph32 : w.graph64, o.check_every, conv);
            st.rz0 = E.rz0_h;
            if (st.rz0 == 0.0) {
                // zero right-hand side: the loop stopped at once and x = 0 is
                // exact (the reference would divide 0/0)
                st.iterations = 0;
                st.converged = 1;
            } else {
                st.iterations = it;
                st.converged = conv ? 1 : 0;
            }
        };
        if (c.precond_kind == 3) run64(float{});
        else run64(double{});
    } else {
        // mixed precision: FP32 SpMV / MAS / Krylov vectors with FP64 dots,
        // outer iterative refinement on the FP64 true residual
        if (!mas && c.precond_kind != 0 && c.precond_kind != 1) throw arg_error("solve: unsupported preconditioner");
        float* rf = w.rf.reserve(N + 1);
        double* r64 = w.r64b.reserve(N + 1);
        double* d64 = w.d64.reserve(N + 1);
        Engine<float, float, float> E{c, N, gN, gS,
                                      rf, w.zf.reserve(N + 1), w.pf.reserve(N + 1),
                                      w.Apf.reserve(N + 1),
                                      d64, {parts, parts + kPart}, parts + 2 * kPart, parts + 3 * kPart,
                                      parts + 4 * kPart, sc, ctrl,
                                      c.spmv_full32 ? nullptr : bsr_values_f32(c),  // the full copy replaces it
                                      mas,
                                      mas ? c.mas.R32.p : nullptr, mas ? c.mas.Z32.p : nullptr,
                                      c.precond_kind == 1 ? w.Dinv.p : nullptr, 1};
        E.G = G;
        if (c.spmv_full32) E.vfull = bsr_full_f32(c);
        fill<double>(x, 0.0, N, s);
        hipLaunchKernelGGL((k_copy_cast<double, double>), dim3(gN), dim3(kThreads), 0, s, N, r64, b);
        SGIPC_LAUNCH_CHECK();
        double* part_o = parts + 5 * kPart;
        // ||b||^2 stays on the device until the first true residual is read
        // with it (the first inner tolerance needs only ||r|| = ||b||)
        constexpr int S_B2 = S_NUM + 1;
        device_norm2_to(c, b, N, part_o, sc, S_B2);
        int total = 0;
        bool conv = false, have = false;  // have: rr, bb read from the device
        double bb = 0.0, rr = 0.0, bn = 0.0, rn = 0.0;  // rr: ||b - A x||^2 of the current x
        st.rz0 = 0.0;
        for (int outer = 0; outer < 50 && total < max_iter; ++outer) {
            if (have && (rn <= o.tol * bn || bn == 0.0)) { conv = true; break; }
            double eta = have ? std::max(0.9 * o.tol * bn / rn, o.inner_tol) : std::max(0.9 * o.tol, o.inner_tol);
            hipLaunchKernelGGL((k_copy_cast<float, double>), dim3(gN), dim3(kThreads), 0, s, N, rf, r64);
            SGIPC_LAUNCH_CHECK();
            E.init(eta, max_iter - total);
            bool c2 = false;
            int it = E.run(w.graph_mixed, o.check_every, c2);
            if (outer == 0) st.rz0 = E.rz0_h;
            if (E.rz0_h == 0.0) { conv = true; break; }  // the residual is zero: x is exact
            total += it;
            hipLaunchKernelGGL(k_axpy, dim3(gN), dim3(kThreads), 0, s, N, x, 1.0, (const double*)d64);
            SGIPC_LAUNCH_CHECK();
            // FP64 true residual r = b - A x
            hipLaunchKernelGGL((k_copy_cast<double, double>), dim3(gN), dim3(kThreads), 0, s, N, r64, b);
            SGIPC_LAUNCH_CHECK();
            spmv64(c, -1.0, x, 1.0, r64);
            device_norm2_to(c, r64, N, part_o, sc, S_NUM);
            read_scalars(c, sc, S_NUM, S_B2, rr, bb);
            have = true;
            rn = std::sqrt(rr);
            bn = std::sqrt(bb);
            st.refinements = outer + 1;
        }
        if (!have) {  // no residual computed (zero right-hand side or no budget): x = 0, r = b
            read_scalars(c, sc, S_B2, S_B2, bb, bb);
            rr = bb;
            rn = bn = std::sqrt(bb);
        }
        if (rn <= o.tol * bn) conv = true;
        st.iterations = total;
        st.converged = conv ? 1 : 0;
        // the refinement loop's last FP64 true residual is that of the final x
        st.rel_residual = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
        return;
    }
    // true residual ||b - A x|| / ||b||, FP64
    double* tmp = w.xw.reserve(N + 1);
    hipLaunchKernelGGL((k_copy_cast<double, double>), dim3(gN), dim3(kThreads), 0, s, N, tmp, b);
    SGIPC_LAUNCH_CHECK();
    spmv64(c, -1.0, x, 1.0, tmp);
    // ||b||^2 is the init's S_BB (r held b); one read for both
    device_norm2_to(c, tmp, N, parts + 5 * kPart, sc, S_NUM);
    double rr = 0.0, bb = 0.0;
    read_scalars(c, sc, S_NUM, S_BB, rr, bb);
    st.rel_residual = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
}

}  // namespace sgipc
