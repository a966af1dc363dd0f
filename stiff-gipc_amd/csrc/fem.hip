// FEM part of the Newton-step assembly: per-tet Stable Neo-Hookean (SNK1)
// gradient and PSD-projected 12x12 Hessian written as 10 upper 3x3 triplets,
// deterministic per-vertex gradient gather, kinetic gradient, mass diagonal
// triplets and the FEM right-hand side.
//
// Reference: S/femEnergy.cu:1719-1857 (_calculate_fem_gradient_hessian),
// :432-459 (SNK1 PK1), :1593-1717 (analytic eigensystem + clamp),
// S/QRSVD.hpp:219-317 (signed 3x3 SVD), S/GIPC.cu:6339-6350 (kinetic),
// :10295-10336 (boundary mask, mass diagonal),
// S/linear_system/subsystem/fem_linear_subsystem.cu:66-83 (RHS).
#include "context.hpp"
#include <algorithm>

namespace sgipc {
namespace {

// ---- 3x3 signed SVD: implicit-shift QR on the bidiagonal form -------------
// Published method: Gast, Jiang, Teran, "Implicit-shifted symmetric QR
// singular value decomposition of 3x3 matrices" (UCLA, 2016).  Matrices are
// row-major double[9]; rotation indices are template parameters so that all
// arrays stay in registers.
struct Rot {
    double c, s;
};

__device__ inline Rot rot_conv(double a, double b) {
    double d = a * a + b * b;
    Rot r{1.0, 0.0};
    if (d != 0.0) {
        double t = rsqrt(d);
        r.c = a * t;
        r.s = -b * t;
    }
    return r;
}
__device__ inline Rot rot_unconv(double a, double b) {
    double d = a * a + b * b;
    Rot r{0.0, 1.0};
    if (d != 0.0) {
        double t = rsqrt(d);
        r.s = a * t;
        r.c = b * t;
    }
    return r;
}
template <int I, int K>
__device__ inline void rows(const Rot& g, double* A) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        double t1 = A[3 * I + j], t2 = A[3 * K + j];
        A[3 * I + j] = g.c * t1 - g.s * t2;
        A[3 * K + j] = g.s * t1 + g.c * t2;
    }
}
template <int I, int K>
__device__ inline void cols(const Rot& g, double* A) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double t1 = A[3 * r + I], t2 = A[3 * r + K];
        A[3 * r + I] = g.c * t1 - g.s * t2;
        A[3 * r + K] = g.s * t1 + g.c * t2;
    }
}

__device__ inline void zero_chase(double* H, double* U, double* V) {
    Rot r1 = rot_conv(H[0], H[3]);
    Rot r2 = (H[3] != 0.0) ? rot_conv(H[0] * H[1] + H[3] * H[4], H[0] * H[2] + H[3] * H[5])
                           : rot_conv(H[1], H[2]);
    rows<0, 1>(r1, H);
    cols<1, 2>(r2, H);
    cols<1, 2>(r2, V);
    Rot r3 = rot_conv(H[4], H[7]);
    rows<1, 2>(r3, H);
    cols<0, 1>(r1, U);
    cols<1, 2>(r3, U);
}

// 2x2 SVD (polar + symmetric Jacobi) of block (T,T); applies to U, V cols T,T+1
template <int T>
__device__ inline void process(const double* B, double* U, double* S, double* V) {
    constexpr int O = (T == 1) ? 0 : 2;
    S[O] = B[4 * O];
    double a00 = B[3 * T + T], a01 = B[3 * T + T + 1], a10 = B[3 * (T + 1) + T], a11 = B[3 * (T + 1) + T + 1];
    double x0 = a00 + a11, x1 = a10 - a01, d = sqrt(x0 * x0 + x1 * x1);
    Rot u{1.0, 0.0};
    if (d != 0.0) { u.c = x0 / d; u.s = -x1 / d; }
    double s00 = u.c * a00 - u.s * a10, s01 = u.c * a01 - u.s * a11, s11 = u.s * a01 + u.c * a11;
    double cs, sn, S0, S1, y2 = s01 * s01;
    if (y2 == 0.0) { cs = 1.0; sn = 0.0; S0 = s00; S1 = s11; }
    else {
        double tau = 0.5 * (s00 - s11), w = sqrt(tau * tau + y2);
        double t = (tau > 0.0) ? s01 / (tau + w) : s01 / (tau - w);
        cs = rsqrt(t * t + 1.0);
        sn = -t * cs;
        double c2 = cs * cs, c2y = 2.0 * cs * sn * s01, s2 = sn * sn;
        S0 = c2 * s00 - c2y + s2 * s11;
        S1 = s2 * s00 + c2y + c2 * s11;
    }
    Rot v;
    if (S0 < S1) { double tt = S0; S0 = S1; S1 = tt; v.c = -sn; v.s = cs; }
    else { v.c = cs; v.s = sn; }
    Rot uu{u.c * v.c - u.s * v.s, u.s * v.c + u.c * v.s};
    S[T] = S0; S[T + 1] = S1;
    cols<T, T + 1>(uu, U);
    cols<T, T + 1>(v, V);
}

template <int J>
__device__ inline void negc(double* M) {
#pragma unroll
    for (int r = 0; r < 3; ++r) M[3 * r + J] = -M[3 * r + J];
}
template <int A, int B>
__device__ inline void swapc(double* M) {
#pragma unroll
    for (int r = 0; r < 3; ++r) { double t = M[3 * r + A]; M[3 * r + A] = M[3 * r + B]; M[3 * r + B] = t; }
}
__device__ inline void swp(double& a, double& b) { double t = a; a = b; b = t; }

__device__ inline void sort0(double* U, double* S, double* V) {
    if (fabs(S[1]) >= fabs(S[2])) {
        if (S[1] < 0) { S[1] = -S[1]; S[2] = -S[2]; negc<1>(U); negc<2>(U); }
        return;
    }
    if (S[2] < 0) { S[1] = -S[1]; S[2] = -S[2]; negc<1>(U); negc<2>(U); }
    swp(S[1], S[2]); swapc<1, 2>(U); swapc<1, 2>(V);
    if (S[1] > S[0]) { swp(S[0], S[1]); swapc<0, 1>(U); swapc<0, 1>(V); }
    else { negc<2>(U); negc<2>(V); }
}
__device__ inline void sort1(double* U, double* S, double* V) {
    if (fabs(S[0]) >= S[1]) {
        if (S[0] < 0) { S[0] = -S[0]; S[2] = -S[2]; negc<0>(U); negc<2>(U); }
        return;
    }
    swp(S[0], S[1]); swapc<0, 1>(U); swapc<0, 1>(V);
    if (fabs(S[1]) < fabs(S[2])) { swp(S[1], S[2]); swapc<1, 2>(U); swapc<1, 2>(V); }
    else { negc<1>(U); negc<1>(V); }
    if (S[1] < 0) { S[1] = -S[1]; S[2] = -S[2]; negc<1>(U); negc<2>(U); }
}

__device__ void svd3(const double* A, double* U, double* S, double* V) {
    double B[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) { B[i] = A[i]; U[i] = (i % 4 == 0) ? 1.0 : 0.0; V[i] = U[i]; }
    {
        Rot r = rot_conv(B[3], B[6]);
        rows<1, 2>(r, B);
        cols<1, 2>(r, U);
        zero_chase(B, U, V);
    }
    const double eta = 2.220446049250313e-16 * 128.0;
    double nrm = 0.5 * sqrt(B[0] * B[0] + B[4] * B[4] + B[8] * B[8] + B[1] * B[1] + B[5] * B[5]);
    double tol = eta * fmax(nrm, 1.0);
    for (int it = 0; it < 100; ++it) {
        double a0 = B[0], a1 = B[4], a2 = B[8], b0 = B[1], b1 = B[5];
        if (!(fabs(a0) > tol && fabs(a1) > tol && fabs(a2) > tol && fabs(b0) > tol && fabs(b1) > tol)) break;
        double t11 = a1 * a1 + b0 * b0, t22 = a2 * a2 + b1 * b1, g1 = a1 * b1;
        double dd = 0.5 * (t11 - t22), bs = g1 * g1;
        double mu = t22 - copysign(bs / (fabs(dd) + sqrt(dd * dd + bs)), dd);
        Rot r = rot_conv(a0 * a0 - mu, a0 * b0);
        cols<0, 1>(r, B);
        cols<0, 1>(r, V);
        zero_chase(B, U, V);
    }
    if (fabs(B[5]) <= tol) {
        process<0>(B, U, S, V); sort0(U, S, V);
    } else if (fabs(B[1]) <= tol) {
        process<1>(B, U, S, V); sort1(U, S, V);
    } else if (fabs(B[4]) <= tol) {
        Rot r1 = rot_unconv(B[5], B[8]);
        rows<1, 2>(r1, B); cols<1, 2>(r1, U);
        process<0>(B, U, S, V); sort0(U, S, V);
    } else if (fabs(B[8]) <= tol) {
        Rot r1 = rot_conv(B[4], B[5]);
        cols<1, 2>(r1, B); cols<1, 2>(r1, V);
        Rot r2 = rot_conv(B[0], B[2]);
        cols<0, 2>(r2, B); cols<0, 2>(r2, V);
        process<0>(B, U, S, V); sort0(U, S, V);
    } else {
        Rot r1 = rot_unconv(B[1], B[4]);
        rows<0, 1>(r1, B); cols<0, 1>(r1, U);
        Rot r2 = rot_unconv(B[2], B[8]);
        rows<0, 2>(r2, B); cols<0, 2>(r2, U);
        process<1>(B, U, S, V); sort1(U, S, V);
    }
}

// symmetric 3x3 eigen-decomposition by cyclic Jacobi (eigenvectors in
// columns of Q, row-major); eigenvalue order is irrelevant to the projection.
template <int P, int Q>
__device__ inline void jrot(double* a, double* E) {
    double apq = a[3 * P + Q];
    if (apq == 0.0) return;
    double theta = (a[3 * Q + Q] - a[3 * P + P]) / (2.0 * apq);
    double t = copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
    double c = rsqrt(t * t + 1.0), s = t * c;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double akp = a[3 * k + P], akq = a[3 * k + Q];
        a[3 * k + P] = c * akp - s * akq;
        a[3 * k + Q] = s * akp + c * akq;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double apk = a[3 * P + k], aqk = a[3 * Q + k];
        a[3 * P + k] = c * apk - s * aqk;
        a[3 * Q + k] = s * apk + c * aqk;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double ekp = E[3 * k + P], ekq = E[3 * k + Q];
        E[3 * k + P] = c * ekp - s * ekq;
        E[3 * k + Q] = s * ekp + c * ekq;
    }
}
__device__ void syev3(double* a, double* w, double* E) {
#pragma unroll
    for (int i = 0; i < 9; ++i) E[i] = (i % 4 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 32; ++sweep) {
        double off = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
        double dg = a[0] * a[0] + a[4] * a[4] + a[8] * a[8];
        if (off <= 1e-36 * dg || off == 0.0) break;
        jrot<0, 1>(a, E);
        jrot<0, 2>(a, E);
        jrot<1, 2>(a, E);
    }
    w[0] = a[0]; w[1] = a[4]; w[2] = a[8];
}

__device__ inline double det3(const double* F) {
    return F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
           F[2] * (F[3] * F[7] - F[4] * F[6]);
}

// Hq (9x9, vec index i+3j) = sum_k max(lambda_k,0) q_k q_k^T, upper triangle
// packed row-wise into 45 entries.
__device__ inline int up9(int a, int b) { return a * 9 - a * (a - 1) / 2 + (b - a); }

__device__ inline void accum_mode(double* Hq, double lam, const double* q, bool project) {
    if (project && lam <= 0.0) return;
#pragma unroll
    for (int a = 0; a < 9; ++a) {
        double la = lam * q[a];
#pragma unroll
        for (int b = a; b < 9; ++b) Hq[up9(a, b)] += la * q[b];
    }
}

// project: clamp negative modes (the PSD Newton Hessian); otherwise the
// exact second derivative (all nine modes), used by adjoint solves
__device__ void snk1_hq(const double* F, double mu, double lam, double* Hq, bool project) {
    double U[9], S[3], V[9];
    svd3(F, U, S, V);
    double J = det3(F);
#pragma unroll
    for (int i = 0; i < 45; ++i) Hq[i] = 0.0;
    const double h = 0.70710678118654752440;
    double sc = lam * (J - 1.0) - mu;
    // twist / flip modes
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const int p = (t == 0) ? 1 : 0, qq = (t == 2) ? 1 : 2;
        double tw[9], fl[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double a = h * U[3 * r + p] * V[3 * c + qq];
                double b = h * U[3 * r + qq] * V[3 * c + p];
                tw[r + 3 * c] = b - a;
                fl[r + 3 * c] = a + b;
            }
        accum_mode(Hq, mu + S[t] * sc, tw, project);
        accum_mode(Hq, mu - S[t] * sc, fl, project);
    }
    // scaling modes: 3x3 block in singular-value space
    double s00 = S[0] * S[0], s11 = S[1] * S[1], s22 = S[2] * S[2];
    double sc2 = lam * (2.0 * J - 1.0) - mu;
    double Am[9] = {mu + lam * s11 * s22, sc2 * S[2], sc2 * S[1],
                    sc2 * S[2], mu + lam * s00 * s22, sc2 * S[0],
                    sc2 * S[1], sc2 * S[0], mu + lam * s00 * s11};
    double w[3], E[9];
    syev3(Am, w, E);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (project && w[k] <= 0.0) continue;
        double q[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                q[r + 3 * c] = U[3 * r + 0] * E[0 + k] * V[3 * c + 0] + U[3 * r + 1] * E[3 + k] * V[3 * c + 1] +
                               U[3 * r + 2] * E[6 + k] * V[3 * c + 2];
        accum_mode(Hq, w[k], q, project);
    }
}

__device__ inline double hq(const double* Hq, int a, int b) { return a <= b ? Hq[up9(a, b)] : Hq[up9(b, a)]; }

__global__ void __launch_bounds__(128) k_fem_snk1(int T, const double* __restrict__ x,
                                                  const uint4* __restrict__ tets,
                                                  const double* __restrict__ DmInv,
                                                  const double* __restrict__ vol,
                                                  const double* __restrict__ mu,
                                                  const double* __restrict__ lam, double dt2,
                                                  const int* __restrict__ btype, int index_offset,
                                                  double* __restrict__ tvals, int* __restrict__ trows,
                                                  int* __restrict__ tcols, double* __restrict__ tet_grad,
                                                  int project) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    uint4 tv = tets[t];
    const unsigned vid[4] = {tv.x, tv.y, tv.z, tv.w};
    double X[12];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) X[3 * a + k] = x[3 * (size_t)vid[a] + k];
    double Dm[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Dm[i] = DmInv[9 * (size_t)t + i];
    // Bm (4x3): dF(i,j)/dx_{a,k} = delta_ik Bm(a,j)
    double Bm[12];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        Bm[j] = -(Dm[j] + Dm[3 + j] + Dm[6 + j]);
        Bm[3 + j] = Dm[j];
        Bm[6 + j] = Dm[3 + j];
        Bm[9 + j] = Dm[6 + j];
    }
    // F(i,j) = sum_a X_a,i Bm(a,j)
    double F[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            F[3 * i + j] = X[i] * Bm[j] + X[3 + i] * Bm[3 + j] + X[6 + i] * Bm[6 + j] + X[9 + i] * Bm[9 + j];
    double m = mu[t], l = lam[t];
    double s = dt2 * vol[t];
    // gradient: P = mu F + lam (J - 1 - mu/lam) cof(F)
    {
        double cof[9];
        cof[0] = F[4] * F[8] - F[5] * F[7]; cof[1] = F[5] * F[6] - F[3] * F[8]; cof[2] = F[3] * F[7] - F[4] * F[6];
        cof[3] = F[7] * F[2] - F[8] * F[1]; cof[4] = F[8] * F[0] - F[6] * F[2]; cof[5] = F[6] * F[1] - F[7] * F[0];
        cof[6] = F[1] * F[5] - F[4] * F[2]; cof[7] = F[2] * F[3] - F[0] * F[5]; cof[8] = F[0] * F[4] - F[1] * F[3];
        double J = det3(F);
        double kk = l * (J - 1.0 - m / l);
        double* g = tet_grad + 12 * (size_t)t;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double acc = 0.0;
#pragma unroll
                for (int j = 0; j < 3; ++j) acc += Bm[3 * a + j] * (m * F[3 * k + j] + kk * cof[3 * k + j]);
                g[3 * a + k] = s * acc;
            }
    }
    double Hq[45];
    snk1_hq(F, m, l, Hq, project != 0);
    bool pin[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) pin[a] = btype ? (btype[vid[a]] != 0) : false;
    size_t base = (size_t)t * 10;
    int kk = 0;
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
#pragma unroll
        for (int jj = ii; jj < 4; ++jj, ++kk) {
            int r = (int)vid[ii] + index_offset, c = (int)vid[jj] + index_offset;
            bool tr = r > c;
            bool zero = pin[ii] || pin[jj];
            double blk[9];  // column-major: blk[p + 3q] = H(ii*3+p, jj*3+q) (or transposed)
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    double acc = 0.0;
#pragma unroll
                    for (int j = 0; j < 3; ++j)
#pragma unroll
                        for (int jp = 0; jp < 3; ++jp)
                            acc += Bm[3 * ii + j] * Bm[3 * jj + jp] * hq(Hq, p + 3 * j, q + 3 * jp);
                    acc = zero ? 0.0 : s * acc;
                    if (tr) blk[q + 3 * p] = acc; else blk[p + 3 * q] = acc;
                }
            double* dst = tvals + 9 * (base + kk);
#pragma unroll
            for (int e = 0; e < 9; ++e) dst[e] = blk[e];
            trows[base + kk] = tr ? c : r;
            tcols[base + kk] = tr ? r : c;
        }
}

// Parameter Jacobian of the SNK1 gradient contracted with an adjoint vector:
// the tet gradient is g_a = dt^2 vol Bm(a,:) P^T with
// P = mu F + (lam (J - 1) - mu) cof(F), linear in (mu, lam), so
//   d(adj . g)/d mu  = dt^2 vol <G, F - cof(F)>
//   d(adj . g)/d lam = dt^2 vol (J - 1) <G, cof(F)>
// with G(k, j) = sum_a adj(v_a, k) Bm(a, j); pinned vertices contribute no
// gradient rows (their right-hand side is zeroed), so their adjoint is
// taken as zero.  No reference counterpart (the reference has no adjoint);
// the gradient it differentiates is _calculate_fem_gradient_hessian
// (StiffGIPC/femEnergy.cu:1719-1857).
__global__ void __launch_bounds__(128) k_fem_param_adjoint(int T, const double* __restrict__ x,
                                                           const uint4* __restrict__ tets,
                                                           const double* __restrict__ DmInv,
                                                           const double* __restrict__ vol, double dt2,
                                                           const int* __restrict__ btype,
                                                           const double* __restrict__ adj,
                                                           double* __restrict__ dmu, double* __restrict__ dlam) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    uint4 tv = tets[t];
    const unsigned vid[4] = {tv.x, tv.y, tv.z, tv.w};
    double X[12], A[12];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        bool pin = btype ? (btype[vid[a]] != 0) : false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            X[3 * a + k] = x[3 * (size_t)vid[a] + k];
            A[3 * a + k] = pin ? 0.0 : adj[3 * (size_t)vid[a] + k];
        }
    }
    double Dm[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Dm[i] = DmInv[9 * (size_t)t + i];
    double Bm[12];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        Bm[j] = -(Dm[j] + Dm[3 + j] + Dm[6 + j]);
        Bm[3 + j] = Dm[j];
        Bm[6 + j] = Dm[3 + j];
        Bm[9 + j] = Dm[6 + j];
    }
    double F[9], G[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            F[3 * i + j] = X[i] * Bm[j] + X[3 + i] * Bm[3 + j] + X[6 + i] * Bm[6 + j] + X[9 + i] * Bm[9 + j];
            G[3 * i + j] = A[i] * Bm[j] + A[3 + i] * Bm[3 + j] + A[6 + i] * Bm[6 + j] + A[9 + i] * Bm[9 + j];
        }
    double cof[9];
    cof[0] = F[4] * F[8] - F[5] * F[7]; cof[1] = F[5] * F[6] - F[3] * F[8]; cof[2] = F[3] * F[7] - F[4] * F[6];
    cof[3] = F[7] * F[2] - F[8] * F[1]; cof[4] = F[8] * F[0] - F[6] * F[2]; cof[5] = F[6] * F[1] - F[7] * F[0];
    cof[6] = F[1] * F[5] - F[4] * F[2]; cof[7] = F[2] * F[3] - F[0] * F[5]; cof[8] = F[0] * F[4] - F[1] * F[3];
    double J = det3(F);
    double gf = 0.0, gc = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        gf += G[i] * F[i];
        gc += G[i] * cof[i];
    }
    double s = dt2 * vol[t];
    dmu[t] = s * (gf - gc);
    dlam[t] = s * (J - 1.0) * gc;
}

// shape[v] += sum over incident tets (ascending tet id) of the tet gradient
__global__ void k_gather_tet_grad(int n, const int* __restrict__ inc_start, const int* __restrict__ inc_list,
                                  const double* __restrict__ tet_grad, double* __restrict__ shape) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    double g0 = shape[3 * (size_t)v], g1 = shape[3 * (size_t)v + 1], g2 = shape[3 * (size_t)v + 2];
    int e = inc_start[v + 1];
    for (int i = inc_start[v]; i < e; ++i) {
        int tc = inc_list[i];
        const double* g = tet_grad + 3 * (size_t)tc;  // 12*t + 3*corner = 3*(4t+corner)
        g0 += g[0]; g1 += g[1]; g2 += g[2];
    }
    shape[3 * (size_t)v] = g0; shape[3 * (size_t)v + 1] = g1; shape[3 * (size_t)v + 2] = g2;
}

__global__ void k_kinetic(int n, const double* __restrict__ x, const double* __restrict__ xt,
                          const double* __restrict__ mass, double* __restrict__ g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    g[i] = (x[i] - xt[i]) * mass[i / 3];
}

__global__ void k_mass_triplets(int n, const double* __restrict__ mass, int off, double* __restrict__ vals,
                                int* __restrict__ rows, int* __restrict__ cols) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double m = mass[i];
#pragma unroll
    for (int k = 0; k < 9; ++k) vals[9 * (size_t)i + k] = (k % 4 == 0) ? m : 0.0;
    rows[i] = i + off;
    cols[i] = i + off;
}

__global__ void k_fem_rhs(int n, const double* __restrict__ fb, const double* __restrict__ sh,
                          const int* __restrict__ btype, double* __restrict__ b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    int v = i / 3;
    double fbv = fb ? fb[i] : 0.0;
    b[i] = (btype && btype[v] != 0) ? 0.0 : fbv + sh[i];
}

}  // namespace

void fem_set_mesh(Ctx& c, int n, int T, const uint32_t* h_tets) {
    if (n < 0 || T < 0) throw arg_error("fem_set_mesh: negative size");
    std::vector<int> cnt((size_t)n + 1, 0);
    for (long long i = 0; i < 4LL * T; ++i) {
        uint32_t v = h_tets[i];
        if ((long long)v >= n) throw arg_error("fem_set_mesh: tet vertex index out of range");
        cnt[v + 1]++;
    }
    for (int v = 0; v < n; ++v) cnt[v + 1] += cnt[v];
    std::vector<int> lst((size_t)4 * T + 1), cur(cnt.begin(), cnt.end() - 1);
    for (int t = 0; t < T; ++t)
        for (int a = 0; a < 4; ++a) lst[cur[h_tets[4 * (size_t)t + a]]++] = 4 * t + a;
    c.fem_n = n;
    c.fem_T = T;
    c.inc_start.reserve((size_t)n + 1);
    c.inc_list.reserve((size_t)4 * T + 1);
    c.tet_grad.reserve((size_t)12 * T + 1);
    SGIPC_HIP(hipMemcpy(c.inc_start.p, cnt.data(), sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice));
    if (T > 0)
        SGIPC_HIP(hipMemcpy(c.inc_list.p, lst.data(), sizeof(int) * 4 * (size_t)T, hipMemcpyHostToDevice));
}

void fem_kinetic_gradient(Ctx& c, int n, const double* x, const double* xt, const double* mass, double* shape) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_kinetic, dim3(div_up(3LL * n, 256)), dim3(256), 0, c.stream, n, x, xt, mass, shape);
    SGIPC_LAUNCH_CHECK();
}

void fem_assemble(Ctx& c, int T, const double* x, const uint32_t* tets, const double* DmInv, const double* vol,
                  const double* mu, const double* lam, double dt, const int* btype, int index_offset,
                  long long triplet_offset, double* shape) {
    if (T != c.fem_T) throw state_error("fem_assemble: tet count differs from sgipc_set_fem_mesh");
    if (!c.t_vals) throw state_error("fem_assemble: no triplet view set");
    if (triplet_offset < 0 || triplet_offset + 10LL * T > c.t_cap)
        throw capacity_error("fem_assemble: triplet capacity exceeded");
    if (T <= 0) return;
    prof_begin(c, PROF_FEM);
    hipLaunchKernelGGL(k_fem_snk1, dim3(div_up(T, 128)), dim3(128), 0, c.stream, T, x, (const uint4*)tets, DmInv,
                       vol, mu, lam, dt * dt, btype, index_offset, c.t_vals + 9 * triplet_offset,
                       c.t_rows + triplet_offset, c.t_cols + triplet_offset, c.tet_grad.p, c.fem_project);
    SGIPC_LAUNCH_CHECK();
    prof_end(c, PROF_FEM);
    hipLaunchKernelGGL(k_gather_tet_grad, dim3(div_up(c.fem_n, 256)), dim3(256), 0, c.stream, c.fem_n,
                       c.inc_start.p, c.inc_list.p, c.tet_grad.p, shape);
    SGIPC_LAUNCH_CHECK();
}

void fem_param_adjoint(Ctx& c, int T, const double* x, const uint32_t* tets, const double* DmInv, const double* vol,
                       double dt, const int* btype, const double* adj, double* dmu, double* dlam) {
    if (T <= 0) return;
    hipLaunchKernelGGL(k_fem_param_adjoint, dim3(div_up(T, 128)), dim3(128), 0, c.stream, T, x, (const uint4*)tets,
                       DmInv, vol, dt * dt, btype, adj, dmu, dlam);
    SGIPC_LAUNCH_CHECK();
}

void fem_mass_triplets(Ctx& c, int n, const double* mass, int index_offset, long long triplet_offset) {
    if (!c.t_vals) throw state_error("mass_triplets: no triplet view set");
    if (triplet_offset < 0 || triplet_offset + n > c.t_cap) throw capacity_error("mass_triplets: capacity exceeded");
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mass_triplets, dim3(div_up(n, 256)), dim3(256), 0, c.stream, n, mass, index_offset,
                       c.t_vals + 9 * triplet_offset, c.t_rows + triplet_offset, c.t_cols + triplet_offset);
    SGIPC_LAUNCH_CHECK();
}

void fem_rhs(Ctx& c, int n, const double* fb, const double* shape, const int* btype, double* b) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fem_rhs, dim3(div_up(3LL * n, 256)), dim3(256), 0, c.stream, n, fb, shape, btype, b);
    SGIPC_LAUNCH_CHECK();
}

}  // namespace sgipc
