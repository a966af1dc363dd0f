// Affine bodies (ABD) coupled to the FEM system: the stiff bodies of the
// hybrid configuration.  Body b has 12 DOF q = (p, a1, a2, a3) and vertex i
// of the body sits at x_i = p + (a1 . xbar_i, a2 . xbar_i, a3 . xbar_i)
// = J_i q.  Vertex space: the nav body vertices first (grouped by body),
// then the FEM vertices.  DOF space: 4 block rows of 3 per body, then one
// block row per FEM vertex (FEM vertex v -> block row v - nav + 4 nb).
//
// Reference: ABDJacobi (StiffGIPC/abd_system/abd_jacobi_matrix.h:7-60),
// shape energy sum_i (|a_i|^2 - 1)^2 + 2 sum_{i<j} (a_i . a_j)^2
// (abd_system/details/abd_energy.inl:9-27), body gradient and Hessian
// ABDSystem::_cal_abd_body_gradient_and_hessian
// (abd_system/abd_system_function/setup_abd_system_gradient_and_hessian.cu:
// 201-345: kinetic M (q - q~), shape term kappa V dt^2 with the 9x9 block
// made PSD; fixed bodies: G = 0, H = M), write_abd_body_hessian (:132-141,
// 10 upper blocks), write_barrier_hessian (:87-128, J_i^T H J_j written as
// 16 blocks, those below the diagonal dropped), the ABD-FEM 4-block
// coupling (:521-580), the barrier gradient J^T g (:348-422, fixed bodies
// skipped), energies (cal_abd_energy.cu:6-140) and the 12x12 block-Jacobi
// preconditioner (linear_system/preconditioner/abd_preconditioner.cu:15-38,
// setup_abd_system_gradient_and_hessian.cu:612-655).
#include "context.hpp"
#include "pcg_dev.hpp"

#include <algorithm>
#include <rocprim/device/device_radix_sort.hpp>

namespace sgipc {
namespace {

constexpr int kAbdEThreads = 256;

// shape energy of one body: value, gradient (9: d/da1, d/da2, d/da3) and
// Hessian (9x9 row-major); g / H may be null
__device__ double shape_eval(const double* a, double* g, double* H) {
    double d[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) d[i][j] = a[3 * i] * a[3 * j] + a[3 * i + 1] * a[3 * j + 1] + a[3 * i + 2] * a[3 * j + 2];
    double e = 0.0;
    for (int i = 0; i < 3; ++i) e += (d[i][i] - 1.0) * (d[i][i] - 1.0);
    e += 2.0 * (d[0][1] * d[0][1] + d[0][2] * d[0][2] + d[1][2] * d[1][2]);
    if (g)
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) {
                double s = 4.0 * (d[i][i] - 1.0) * a[3 * i + k];
                for (int j = 0; j < 3; ++j)
                    if (j != i) s += 4.0 * d[i][j] * a[3 * j + k];
                g[3 * i + k] = s;
            }
    // d2/dai dai = 4 (ai.ai - 1) I + 8 ai ai^T + 4 sum_{m != i} am am^T
    // d2/dai daj = 4 (ai.aj) I + 4 aj ai^T                      (i != j)
    if (H)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l) {
                        double h;
                        if (i == j) {
                            h = (k == l ? 4.0 * (d[i][i] - 1.0) : 0.0) + 8.0 * a[3 * i + k] * a[3 * i + l];
                            for (int m = 0; m < 3; ++m)
                                if (m != i) h += 4.0 * a[3 * m + k] * a[3 * m + l];
                        } else {
                            h = (k == l ? 4.0 * d[i][j] : 0.0) + 4.0 * a[3 * j + k] * a[3 * i + l];
                        }
                        H[9 * (3 * i + k) + 3 * j + l] = h;
                    }
    return e;
}

// x_i = J_i q (also the vertex displacement J_i dq, the map being linear)
__global__ void k_abd_positions(int nav, const int* __restrict__ body_of, const double* __restrict__ xbar,
                                const double* __restrict__ q, double* __restrict__ x) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nav) return;
    const double* Q = q + 12 * (size_t)body_of[i];
    const double b0 = xbar[3 * (size_t)i], b1 = xbar[3 * (size_t)i + 1], b2 = xbar[3 * (size_t)i + 2];
#pragma unroll
    for (int k = 0; k < 3; ++k) x[3 * (size_t)i + k] = Q[k] + Q[3 + 3 * k] * b0 + Q[4 + 3 * k] * b1 + Q[5 + 3 * k] * b2;
}

// Body gradient (overwrites grad_q) and the 10 upper 3x3 blocks
// (column-major) of the body Hessian at block rows/cols 4b .. 4b+3; one
// 64-lane workgroup per body.  The 9x9 shape block is made PSD by cyclic
// Jacobi rotations (fixed (p, q) order; at most 30 sweeps, stopping when the
// off-diagonal mass is below 1e-30 of the total), negative eigenvalues
// clamped to zero, H = V max(w, 0) V^T; each rotation updates the two
// columns, then the two rows and V's two columns, one lane per entry.
constexpr int kBodyThreads = 64;
__global__ void __launch_bounds__(kBodyThreads) k_abd_body(int nb, const double* __restrict__ q,
                                                           const double* __restrict__ qt,
                                                           const double* __restrict__ M,
                                                           const double* __restrict__ vol,
                                                           const int* __restrict__ fixed, double kdt2,
                                                           double* __restrict__ tv, int* __restrict__ tr,
                                                           int* __restrict__ tc, double* __restrict__ grad_q) {
    __shared__ double A[9][9], V[9][9], H[9][9];
    __shared__ double dq[12], red[2][kBodyThreads];
    __shared__ int flag;
    const int b = blockIdx.x, t = threadIdx.x;
    if (b >= nb) return;
    const double* Q = q + 12 * (size_t)b;
    const double* Mb = M + 144 * (size_t)b;
    const bool fix = fixed[b] != 0;
    const double w = kdt2 * vol[b];
    if (t < 12) dq[t] = Q[t] - qt[12 * (size_t)b + t];
    // shape Hessian entries, one per lane (i, j: affine rows; k, l: components)
    double a[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) a[e] = Q[3 + e];
    for (int e = t; e < 81; e += kBodyThreads) {
        const int r = e / 9, c = e % 9;
        const int i = r / 3, k = r % 3, j = c / 3, l = c % 3;
        double dii = a[3 * i] * a[3 * i] + a[3 * i + 1] * a[3 * i + 1] + a[3 * i + 2] * a[3 * i + 2];
        double dij = a[3 * i] * a[3 * j] + a[3 * i + 1] * a[3 * j + 1] + a[3 * i + 2] * a[3 * j + 2];
        double h;
        if (i == j) {
            h = (k == l ? 4.0 * (dii - 1.0) : 0.0) + 8.0 * a[3 * i + k] * a[3 * i + l];
            for (int m = 0; m < 3; ++m)
                if (m != i) h += 4.0 * a[3 * m + k] * a[3 * m + l];
        } else {
            h = (k == l ? 4.0 * dij : 0.0) + 4.0 * a[3 * j + k] * a[3 * i + l];
        }
        A[r][c] = fix ? 0.0 : w * h;
        V[r][c] = (r == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (!fix) {
        for (int sweep = 0; sweep < 30; ++sweep) {
            double o = 0.0, tt = 0.0;
            for (int e = t; e < 81; e += kBodyThreads) {
                const double v2 = A[e / 9][e % 9] * A[e / 9][e % 9];
                tt += v2;
                if (e / 9 != e % 9) o += v2;
            }
            red[0][t] = o; red[1][t] = tt;
            __syncthreads();
            if (t == 0) {
                double so = 0.0, st = 0.0;
                for (int u = 0; u < kBodyThreads; ++u) { so += red[0][u]; st += red[1][u]; }
                flag = so <= 1e-30 * st;
            }
            __syncthreads();
            if (flag) break;
            for (int pp = 0; pp < 8; ++pp)
                for (int qq = pp + 1; qq < 9; ++qq) {
                    const double apq = A[pp][qq];  // every lane reads the same value
                    if (apq == 0.0) continue;      // uniform branch
                    const double th = (A[qq][qq] - A[pp][pp]) / (2.0 * apq);
                    const double tn = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                    const double cs = 1.0 / sqrt(tn * tn + 1.0), sn = tn * cs;
                    __syncthreads();               // all lanes read A before it changes
                    if (t < 9) {                   // A <- A R (columns pp, qq)
                        const double akp = A[t][pp], akq = A[t][qq];
                        A[t][pp] = cs * akp - sn * akq;
                        A[t][qq] = sn * akp + cs * akq;
                    }
                    __syncthreads();
                    if (t < 9) {                   // A <- R^T A (rows pp, qq)
                        const double apk = A[pp][t], aqk = A[qq][t];
                        A[pp][t] = cs * apk - sn * aqk;
                        A[qq][t] = sn * apk + cs * aqk;
                    } else if (t < 18) {           // V <- V R
                        const int k = t - 9;
                        const double vkp = V[k][pp], vkq = V[k][qq];
                        V[k][pp] = cs * vkp - sn * vkq;
                        V[k][qq] = sn * vkp + cs * vkq;
                    }
                    __syncthreads();
                }
        }
        for (int e = t; e < 81; e += kBodyThreads) {
            const int i = e / 9, j = e % 9;
            double acc = 0.0;
            for (int k = 0; k < 9; ++k) acc += V[i][k] * fmax(A[k][k], 0.0) * V[j][k];
            H[i][j] = acc;
        }
    } else {
        for (int e = t; e < 81; e += kBodyThreads) H[e / 9][e % 9] = 0.0;
    }
    __syncthreads();
    // gradient: M dq plus the shape term (free bodies), 0 for fixed ones
    if (t < 12) {
        double g = 0.0;
        if (!fix) {
            for (int k = 0; k < 12; ++k) g += Mb[12 * t + k] * dq[k];
            if (t >= 3) {
                const int i = (t - 3) / 3, k = (t - 3) % 3;
                const double dii = a[3 * i] * a[3 * i] + a[3 * i + 1] * a[3 * i + 1] + a[3 * i + 2] * a[3 * i + 2];
                double sg = 4.0 * (dii - 1.0) * a[3 * i + k];
                for (int j = 0; j < 3; ++j)
                    if (j != i)
                        sg += 4.0 * (a[3 * i] * a[3 * j] + a[3 * i + 1] * a[3 * j + 1] + a[3 * i + 2] * a[3 * j + 2]) *
                              a[3 * j + k];
                g += w * sg;
            }
        }
        grad_q[12 * (size_t)b + t] = g;
    }
    // 10 upper blocks x 9 entries
    for (int e = t; e < 90; e += kBodyThreads) {
        const int blk = e / 9, f = e % 9, p = f % 3, cc = f / 3;
        int bi = 0, rem = blk;
        while (rem >= 4 - bi) { rem -= 4 - bi; ++bi; }
        const int bj = bi + rem;
        const int r = 3 * bi + p, c = 3 * bj + cc;
        double h = Mb[12 * r + c];
        if (r >= 3 && c >= 3) h += H[r - 3][c - 3];
        tv[9 * ((size_t)10 * b + blk) + f] = h;
        if (f == 0) {
            tr[10 * (size_t)b + blk] = 4 * b + bi;
            tc[10 * (size_t)b + blk] = 4 * b + bj;
        }
    }
}

// Block (k, l) of J_i^T H J_j for a 3x3 column-major H: entry (a, c) =
// J_i(u, 3k + a) H(u, v) J_j(v, 3l + c) with J(u, 3k + a) = delta_ua for
// k = 0 and delta_{u, k-1} xbar_a otherwise (one nonzero per column).
__device__ inline void jthj_block(const double* xi, const double* xj, const double* H, int k, int l, double* out) {
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) {
            const double ja = (k == 0) ? 1.0 : xi[a];
            const double jc = (l == 0) ? 1.0 : xj[c];
            const int u = (k == 0) ? a : k - 1;
            const int v = (l == 0) ? c : l - 1;
            out[a + 3 * c] = ja * H[u + 3 * v] * jc;
        }
}

// DOF-space blocks one vertex-space block expands into (0 for an all-zero
// block or one touching a fixed body)
__device__ inline int map_count(int r, int c, const double* H, int nav, const int* body_of, const int* fixed) {
    bool nz = false;
    for (int e = 0; e < 9; ++e) nz |= H[e] != 0.0;
    if (!nz) return 0;
    const bool ra = r < nav, ca = c < nav;
    if (!ra && !ca) return 1;
    if ((ra && fixed[body_of[r]]) || (ca && fixed[body_of[c]])) return 0;
    if (ra != ca) return 4;
    return body_of[r] == body_of[c] ? 10 : 16;
}

// Vertex-space contact triplets -> DOF space. PASS 0 writes the counts,
// PASS 1 the blocks at the scanned offsets.  The input holds each vertex
// pair once (upper, row <= col); the reference stores the contact blocks in
// full (both (i, j) and (j, i); SymGH is off,
// linear_system/linear_system/global_matrix.h:6) and keeps, per entry, the
// mapped blocks on or above the diagonal.  For a same-body pair i != j that
// is the upper part of X + X^T (X = J_i^T H J_j; the (j, i) entry maps to
// X^T), written here from the single upper entry: the same matrix.
template <int PASS>
__global__ void k_abd_map(long long m, const double* __restrict__ sv, const int* __restrict__ sr,
                          const int* __restrict__ sc, int nav, int nbody, const int* __restrict__ body_of,
                          const int* __restrict__ fixed, const double* __restrict__ xbar, int* __restrict__ cnt, const int* __restrict__ off, double* __restrict__ dv,
                          int* __restrict__ dr, int* __restrict__ dc) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    int r = sr[t], c = sc[t];
    double H[9];
    for (int e = 0; e < 9; ++e) H[e] = sv[9 * t + e];
    const int n = map_count(r, c, H, nav, body_of, fixed);
    if (!PASS) {
        cnt[t] = n;
        return;
    }
    if (n == 0) return;
    if (r > c) {  // row <= col in vertex ids: transpose the block
        int tmp = r; r = c; c = tmp;
        double T[9];
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) T[p + 3 * q] = H[q + 3 * p];
        for (int e = 0; e < 9; ++e) H[e] = T[e];
    }
    const int shift = 4 * nbody - nav;
    long long o = off[t];
    auto put = [&](int row, int col, const double* B) {
        for (int e = 0; e < 9; ++e) dv[9 * o + e] = B[e];
        dr[o] = row;
        dc[o] = col;
        ++o;
    };
    if (r >= nav) {
        put(r + shift, c + shift, H);
        return;
    }
    const double* xi = xbar + 3 * (size_t)r;
    const int bi = body_of[r];
    if (c >= nav) {  // body rows 4bi + k, FEM column: the 4 blocks of J_i^T H
        for (int k = 0; k < 4; ++k) {
            double B[9];
            for (int cc = 0; cc < 3; ++cc)
                for (int a = 0; a < 3; ++a) {
                    const double ja = (k == 0) ? 1.0 : xi[a];
                    const int u = (k == 0) ? a : k - 1;
                    B[a + 3 * cc] = ja * H[u + 3 * cc];
                }
            put(4 * bi + k, c + shift, B);
        }
        return;
    }
    const double* xj = xbar + 3 * (size_t)c;
    const int bj = body_of[c];
    if (bi != bj) {  // bi < bj (bodies ordered by vertex id): all 16 blocks
        for (int k = 0; k < 4; ++k)
            for (int l = 0; l < 4; ++l) {
                double B[9];
                jthj_block(xi, xj, H, k, l, B);
                put(4 * bi + k, 4 * bj + l, B);
            }
        return;
    }
    for (int k = 0; k < 4; ++k)
        for (int l = k; l < 4; ++l) {
            double B[9];
            jthj_block(xi, xj, H, k, l, B);
            if (r != c) {
                double Bt[9];
                jthj_block(xi, xj, H, l, k, Bt);  // block (l, k) of X, transposed into (k, l)
                for (int p = 0; p < 3; ++p)
                    for (int q = 0; q < 3; ++q) B[p + 3 * q] += Bt[q + 3 * p];
            }
            put(4 * bi + k, 4 * bi + l, B);
        }
}

// Duplicate vertex pairs of the contact blocks are merged before the mapping
// (the reference converts its body-body contact blocks first,
// setup_abd_system_gradient_and_hessian.cu:430-440): one key per source
// block, (min, max) vertex pair, or the sentinel for blocks the mapping
// drops (all zero, or touching a fixed body).
constexpr unsigned long long kNoPair = ~0ull;

__global__ void k_abd_pair_keys(long long m, const double* __restrict__ sv, const int* __restrict__ sr,
                                const int* __restrict__ sc, int nav, const int* __restrict__ body_of,
                                const int* __restrict__ fixed, unsigned long long* __restrict__ key,
                                int* __restrict__ idx) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int r = sr[t], c = sc[t];
    double H[9];
    for (int e = 0; e < 9; ++e) H[e] = sv[9 * t + e];
    const bool keep = map_count(r, c, H, nav, body_of, fixed) > 0;
    const unsigned lo = (unsigned)min(r, c), hi = (unsigned)max(r, c);
    key[t] = keep ? (((unsigned long long)lo << 32) | hi) : kNoPair;
    idx[t] = (int)t;
}

__global__ void k_abd_run_heads(long long m, const unsigned long long* __restrict__ key, int* __restrict__ head) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    head[i] = (key[i] != kNoPair && (i == 0 || key[i - 1] != key[i])) ? 1 : 0;
}

// each run head sums its run (ascending source index; blocks stored below
// the diagonal transposed) into one upper vertex-space block
__global__ void k_abd_run_sum(long long m, const unsigned long long* __restrict__ key, const int* __restrict__ idx,
                              const int* __restrict__ head, const int* __restrict__ pos,
                              const double* __restrict__ sv, const int* __restrict__ sr, double* __restrict__ uv,
                              int* __restrict__ ur, int* __restrict__ uc) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || !head[i]) return;
    const unsigned long long k = key[i];
    const int r = (int)(k >> 32), c = (int)(k & 0xffffffffu);
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (long long j = i; j < m && key[j] == k; ++j) {
        const int t = idx[j];
        const double* H = sv + 9 * (size_t)t;
        if (sr[t] == r) {
            for (int e = 0; e < 9; ++e) acc[e] += H[e];
        } else {  // stored at (c, r): add its transpose
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) acc[p + 3 * q] += H[q + 3 * p];
        }
    }
    const int o = pos[i];
    for (int e = 0; e < 9; ++e) uv[9 * (size_t)o + e] = acc[e];
    ur[o] = r;
    uc[o] = c;
}

// grad_q[b] += sum over the body's vertices, in vertex order, of J_i^T g_i
// (fixed bodies unchanged)
__global__ void k_abd_grad(int nb, const int* __restrict__ vstart, const int* __restrict__ fixed,
                           const double* __restrict__ xbar, const double* __restrict__ g, double* __restrict__ grad_q) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb || fixed[b]) return;
    double s[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = vstart[b]; i < vstart[b + 1]; ++i) {
        const double g0 = g[3 * (size_t)i], g1 = g[3 * (size_t)i + 1], g2 = g[3 * (size_t)i + 2];
        const double* xb = xbar + 3 * (size_t)i;
        s[0] += g0; s[1] += g1; s[2] += g2;
        for (int a = 0; a < 3; ++a) {
            s[3 + a] += xb[a] * g0;
            s[6 + a] += xb[a] * g1;
            s[9 + a] += xb[a] * g2;
        }
    }
    for (int k = 0; k < 12; ++k) grad_q[12 * (size_t)b + k] += s[k];
}

// per-body energy 0.5 dq^T M dq (0 for fixed bodies) + kappa V dt^2 S(q),
// one partial per workgroup
__global__ void __launch_bounds__(kAbdEThreads) k_abd_energy(int nb, const double* __restrict__ q,
                                                             const double* __restrict__ qt,
                                                             const double* __restrict__ M,
                                                             const double* __restrict__ vol,
                                                             const int* __restrict__ fixed, double kdt2,
                                                             double* __restrict__ part) {
    __shared__ double sh[kAbdEThreads / 64];
    double e = 0.0;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        const double* Q = q + 12 * (size_t)b;
        if (!fixed[b]) {
            const double* Mb = M + 144 * (size_t)b;
            double dq[12];
            for (int k = 0; k < 12; ++k) dq[k] = Q[k] - qt[12 * (size_t)b + k];
            double kin = 0.0;
            for (int r = 0; r < 12; ++r) {
                double s = 0.0;
                for (int k = 0; k < 12; ++k) s += Mb[12 * r + k] * dq[k];
                kin += dq[r] * s;
            }
            e += 0.5 * kin;
        }
        e += kdt2 * vol[b] * shape_eval(Q + 3, nullptr, nullptr);
    }
    write_partial(e, part, sh);
}

// 12x12 block-Jacobi inverses from the converted BCOO (block rows 4b .. 4b+3;
// upper blocks, diagonal blocks stored whole; row r holds entries
// ustart[r] .. ustart[r + 1], the last row ending at K); one 64-lane
// workgroup per body: Gauss-Jordan with partial pivoting (first largest
// pivot; a zero pivot leaves its row scaled by 0, i.e. z = 0 along it)
constexpr int kJacThreads = 64;
__global__ void __launch_bounds__(kJacThreads) k_abd_jacobi(int nb, int nbr, int K, const int* __restrict__ ustart,
                                                            const int* __restrict__ cols,
                                                            const double* __restrict__ vals,
                                                            double* __restrict__ inv) {
    __shared__ double A[12][25];  // [A | I], one pad column
    __shared__ double fac[12];
    __shared__ int piv_s;
    const int b = blockIdx.x, t = threadIdx.x;
    if (b >= nb) return;
    for (int e = t; e < 12 * 24; e += kJacThreads) A[e / 24][e % 24] = (e % 24 >= 12) ? ((e / 24 == e % 24 - 12) ? 1.0 : 0.0) : 0.0;
    __syncthreads();
    for (int k = 0; k < 4; ++k) {
        const int row = 4 * b + k;
        const int end = row + 1 < nbr ? ustart[row + 1] : K;
        for (int e = ustart[row] + t; e < end; e += kJacThreads) {  // distinct blocks: no write conflicts
            const int l = cols[e] - 4 * b;
            if (l < k || l > 3) continue;
            const double* B = vals + 9 * (size_t)e;
            for (int p = 0; p < 3; ++p)
                for (int qq = 0; qq < 3; ++qq) {
                    A[3 * k + p][3 * l + qq] = B[p + 3 * qq];
                    if (l != k) A[3 * l + qq][3 * k + p] = B[p + 3 * qq];
                }
        }
    }
    __syncthreads();
    for (int c = 0; c < 12; ++c) {
        if (t == 0) {
            int piv = c;
            for (int r = c + 1; r < 12; ++r)
                if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
            piv_s = piv;
        }
        __syncthreads();
        const int piv = piv_s;
        if (piv != c && t < 24) {
            const double tmp = A[c][t];
            A[c][t] = A[piv][t];
            A[piv][t] = tmp;
        }
        __syncthreads();
        const double d = A[c][c];
        const double id = d != 0.0 ? 1.0 / d : 0.0;
        __syncthreads();
        if (t < 24) A[c][t] *= id;
        __syncthreads();
        if (t < 12) fac[t] = A[t][c];
        __syncthreads();
        for (int e = t; e < 12 * 24; e += kJacThreads) {
            const int r = e / 24, k = e % 24;
            if (r != c) A[r][k] -= fac[r] * A[c][k];
        }
        __syncthreads();
    }
    for (int e = t; e < 144; e += kJacThreads) inv[144 * (size_t)b + e] = A[e / 12][12 + e % 12];
}

// z = P_b^-1 r on the body rows (standalone apply)
__global__ void k_abd_apply(int nb, const double* __restrict__ inv, const double* __restrict__ r, double* __restrict__ z) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 12LL * nb) return;
    const long long b = t / 12;
    const double* P = inv + 144 * b + 12 * (t % 12);
    double s = 0.0;
    for (int k = 0; k < 12; ++k) s += P[k] * r[12 * b + k];
    z[t] = s;
}

const AbdState& abd_ready(const Ctx& c, const char* what) {
    if (!c.abd.ready) throw state_error(std::string(what) + ": no affine bodies set");
    return c.abd;
}

}  // namespace

void abd_setup(Ctx& c, int nb, int nav, const int* vstart, const double* xbar, const double* mass12,
               const double* vol, const int* fixed) {
    if (nb < 0 || nav < 0) throw arg_error("abd_setup: negative size");
    if (nb > 0 && (!vstart || !xbar || !mass12 || !vol)) throw arg_error("abd_setup: null array");
    AbdState& a = c.abd;
    std::vector<int> body_of((size_t)nav + 1, 0), fx((size_t)nb + 1, 0);
    if (nb > 0) {
        if (vstart[0] != 0 || vstart[nb] != nav) throw arg_error("abd_setup: body vertex ranges must cover 0 .. nav");
        for (int b = 0; b < nb; ++b) {
            if (vstart[b + 1] <= vstart[b]) throw arg_error("abd_setup: every body needs vertices (ascending ranges)");
            for (int i = vstart[b]; i < vstart[b + 1]; ++i) body_of[i] = b;
            fx[b] = fixed ? (fixed[b] != 0) : 0;
        }
    } else if (nav != 0) {
        throw arg_error("abd_setup: body vertices without bodies");
    }
    a.nb = nb;
    a.nav = nav;
    a.ready = false;
    a.inv_ready = false;
    a.vstart.reserve((size_t)nb + 1);
    a.body_of.reserve((size_t)nav + 1);
    a.fixed.reserve((size_t)nb + 1);
    a.xbar.reserve(3 * (size_t)nav + 1);
    a.mass.reserve(144 * (size_t)nb + 1);
    a.vol.reserve((size_t)nb + 1);
    a.inv.reserve(144 * (size_t)nb + 1);
    if (nb > 0) {
        SGIPC_HIP(hipMemcpy(a.vstart.p, vstart, sizeof(int) * (nb + 1), hipMemcpyHostToDevice));
        SGIPC_HIP(hipMemcpy(a.body_of.p, body_of.data(), sizeof(int) * nav, hipMemcpyHostToDevice));
        SGIPC_HIP(hipMemcpy(a.fixed.p, fx.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
        SGIPC_HIP(hipMemcpy(a.xbar.p, xbar, sizeof(double) * 3 * nav, hipMemcpyHostToDevice));
        SGIPC_HIP(hipMemcpy(a.mass.p, mass12, sizeof(double) * 144 * nb, hipMemcpyHostToDevice));
        SGIPC_HIP(hipMemcpy(a.vol.p, vol, sizeof(double) * nb, hipMemcpyHostToDevice));
    }
    a.ready = nb > 0;
}

void abd_positions(Ctx& c, const double* q, double* x) {
    const AbdState& a = abd_ready(c, "abd_positions");
    hipLaunchKernelGGL(k_abd_positions, dim3(div_up(a.nav, 256)), dim3(256), 0, c.stream, a.nav,
                       (const int*)a.body_of.p, (const double*)a.xbar.p, q, x);
    SGIPC_LAUNCH_CHECK();
}

void abd_assemble(Ctx& c, const double* q, const double* qt, double kappa, double dt, long long toff, double* grad_q) {
    const AbdState& a = abd_ready(c, "abd_assemble");
    if (!c.t_vals) throw state_error("abd_assemble: no triplet view set");
    if (toff < 0 || toff + 10LL * a.nb > c.t_cap) throw capacity_error("abd_assemble: triplet capacity exceeded");
    hipLaunchKernelGGL(k_abd_body, dim3(a.nb), dim3(kBodyThreads), 0, c.stream, a.nb, q, qt,
                       (const double*)a.mass.p, (const double*)a.vol.p, (const int*)a.fixed.p, kappa * dt * dt,
                       c.t_vals + 9 * toff, c.t_rows + toff, c.t_cols + toff, grad_q);
    SGIPC_LAUNCH_CHECK();
}

void abd_contact_gradient(Ctx& c, const double* g, double* grad_q) {
    const AbdState& a = abd_ready(c, "abd_contact_gradient");
    hipLaunchKernelGGL(k_abd_grad, dim3(div_up(a.nb, 64)), dim3(64), 0, c.stream, a.nb, (const int*)a.vstart.p,
                       (const int*)a.fixed.p, (const double*)a.xbar.p, g, grad_q);
    SGIPC_LAUNCH_CHECK();
}

long long abd_map_contacts(Ctx& c, long long src, long long m, long long dst) {
    abd_ready(c, "abd_map_contacts");
    AbdState& a = c.abd;
    if (!c.t_vals) throw state_error("abd_map_contacts: no triplet view set");
    if (src < 0 || m < 0 || src + m > c.t_cap || dst < 0) throw arg_error("abd_map_contacts: bad triplet ranges");
    if (m == 0) return 0;
    if (m > 0x7fffffffLL / 16) throw arg_error("abd_map_contacts: too many triplets");
    hipStream_t s = c.stream;
    const double* sv = c.t_vals + 9 * src;
    const int* sr = c.t_rows + src;
    const int* sc = c.t_cols + src;
    // 1. merge duplicate vertex pairs: stable radix sort of the pair keys,
    //    run heads, fixed-order run sums
    unsigned long long* key = a.key.reserve(2 * (size_t)m + 2);
    int* idx = a.idx.reserve(2 * (size_t)m + 2);
    int* head = a.cnt.reserve((size_t)m + 1);
    int* pos = a.off.reserve((size_t)m + 1);
    int* total_dev = a.total.reserve(2);
    hipLaunchKernelGGL(k_abd_pair_keys, dim3(div_up(m, 256)), dim3(256), 0, s, m, sv, sr, sc, a.nav,
                       (const int*)a.body_of.p, (const int*)a.fixed.p, key, idx);
    SGIPC_LAUNCH_CHECK();
    if (c.sort_impl == SORT_OWN) {
        radix_sort_pairs(c, key, key + m, idx, 0, idx + m, m, 64, s);
    } else {  // rocPRIM as the cross-check
        size_t bytes = 0;
        SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, key, key + m, idx, idx + m, (size_t)m, 0, 64, s));
        char* tmp = a.sort_tmp.reserve(bytes + 1);
        SGIPC_HIP(rocprim::radix_sort_pairs((void*)tmp, bytes, key, key + m, idx, idx + m, (size_t)m, 0, 64, s));
    }
    const unsigned long long* skey = key + m;
    const int* sidx = idx + m;
    hipLaunchKernelGGL(k_abd_run_heads, dim3(div_up(m, 256)), dim3(256), 0, s, m, skey, head);
    SGIPC_LAUNCH_CHECK();
    exclusive_scan(c, head, pos, (int)m, total_dev, s);
    int nu = 0;
    SGIPC_HIP(hipMemcpyAsync(&nu, total_dev, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    if (nu == 0) return 0;
    double* uv = a.uval.reserve(9 * (size_t)nu);
    int* ur = a.urow.reserve((size_t)nu);
    int* uc = a.ucol.reserve((size_t)nu);
    hipLaunchKernelGGL(k_abd_run_sum, dim3(div_up(m, 256)), dim3(256), 0, s, m, skey, sidx, (const int*)head,
                       (const int*)pos, sv, sr, uv, ur, uc);
    SGIPC_LAUNCH_CHECK();
    // 2. map the merged blocks: counts, scan, blocks
    int* cnt = a.cnt.reserve((size_t)m + 1);   // head flags no longer needed
    int* off = a.off.reserve((size_t)m + 1);
    hipLaunchKernelGGL(k_abd_map<0>, dim3(div_up(nu, 256)), dim3(256), 0, s, (long long)nu, (const double*)uv,
                       (const int*)ur, (const int*)uc, a.nav, a.nb, (const int*)a.body_of.p, (const int*)a.fixed.p,
                       (const double*)a.xbar.p, cnt, (const int*)nullptr, (double*)nullptr, (int*)nullptr,
                       (int*)nullptr);
    SGIPC_LAUNCH_CHECK();
    exclusive_scan(c, cnt, off, nu, total_dev, s);
    int total = 0;
    SGIPC_HIP(hipMemcpyAsync(&total, total_dev, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    if (dst + total > c.t_cap) throw capacity_error("abd_map_contacts: triplet capacity exceeded");
    if (dst < src + m && src < dst + total) throw arg_error("abd_map_contacts: output range overlaps the input range");
    if (total == 0) return 0;
    hipLaunchKernelGGL(k_abd_map<1>, dim3(div_up(nu, 256)), dim3(256), 0, s, (long long)nu, (const double*)uv,
                       (const int*)ur, (const int*)uc, a.nav, a.nb, (const int*)a.body_of.p, (const int*)a.fixed.p,
                       (const double*)a.xbar.p, cnt, (const int*)off, c.t_vals + 9 * dst, c.t_rows + dst,
                       c.t_cols + dst);
    SGIPC_LAUNCH_CHECK();
    return total;
}

double abd_energy(Ctx& c, const double* q, const double* qt, double kappa, double dt) {
    const AbdState& a = abd_ready(c, "abd_energy");
    const int g = energy_grid(a.nb);
    double* part = energy_partials(c);
    hipLaunchKernelGGL(k_abd_energy, dim3(g), dim3(kAbdEThreads), 0, c.stream, a.nb, q, qt, (const double*)a.mass.p,
                       (const double*)a.vol.p, (const int*)a.fixed.p, kappa * dt * dt, part);
    SGIPC_LAUNCH_CHECK();
    return energy_finish(c, part, g);
}

void abd_jacobi_build(Ctx& c) {
    AbdState& a = c.abd;
    a.inv_ready = false;
    if (!a.ready) return;
    if (!c.have_bcoo) throw state_error("abd preconditioner: convert first");
    if (c.nb < 4 * a.nb) throw state_error("abd preconditioner: the system has fewer block rows than the bodies");
    hipLaunchKernelGGL(k_abd_jacobi, dim3(a.nb), dim3(kJacThreads), 0, c.stream, a.nb, c.nb, c.K, (const int*)c.uniq_start.p,
                       (const int*)c.t_cols, (const double*)c.t_vals, a.inv.p);
    SGIPC_LAUNCH_CHECK();
    a.inv_ready = true;
}

void abd_precond_apply(Ctx& c, const double* r, double* z) {
    const AbdState& a = c.abd;
    if (!a.inv_ready) return;
    hipLaunchKernelGGL(k_abd_apply, dim3(div_up(12LL * a.nb, 256)), dim3(256), 0, c.stream, a.nb,
                       (const double*)a.inv.p, r, z);
    SGIPC_LAUNCH_CHECK();
}

}  // namespace sgipc
