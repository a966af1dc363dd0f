// IPC barrier gradient + Gauss-Newton Hessian triplets for contact pairs, and
// the narrow-phase classification that produces the int4 pair encoding.
//
// Reference: S/GIPC.cu:3106-4943 (_calBarrierGradientAndHessian, RANK 2,
// NEWF), write_triplet :80-137, fem-fem contact remap :10183-10214, energy
// :701-957; distances S/mlbvh.cu:229-257; classification _dType_PT
// :288-447, _dType_EE :449-532 and the encoders :535-1060.
//
// One thread per pair.  The distance gradient g_d is analytic (the
// reference's generated dF/dx has a single non-zero column, ds/dx with
// s = sqrt(d / dhat)); the non-mollified Hessian is lambda0 g g^T, the
// mollified one a PSD-projected 2x2 Gauss-Newton form in (c, s).
#include "context.hpp"
#include <rocprim/device/device_radix_sort.hpp>
#include "pcg_dev.hpp"

namespace sgipc {
namespace {

struct V3 {
    double x, y, z;
};
__device__ inline V3 ld(const double* p, int i) { return {p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]}; }
__device__ inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline void put(double* g, int slot, V3 v) { g[3 * slot] = v.x; g[3 * slot + 1] = v.y; g[3 * slot + 2] = v.z; }
// element s (0..3, known only at run time) of a four-element register array,
// by selects, so the array stays in registers (an indexed access would put it
// in scratch memory)
__device__ inline V3 pick4(const V3 (&a)[4], int s) {
    V3 r = a[0];
    r = s == 1 ? a[1] : r;
    r = s == 2 ? a[2] : r;
    r = s == 3 ? a[3] : r;
    return r;
}
__device__ inline int pick4(const int (&a)[4], int s) {
    int r = a[0];
    r = s == 1 ? a[1] : r;
    r = s == 2 ? a[2] : r;
    r = s == 3 ? a[3] : r;
    return r;
}
// g[3 s .. 3 s + 2] += v for a run-time slot s of a register-resident g
__device__ inline void add_at(double (&g)[12], int s, V3 v) {
    g[0] += s == 0 ? v.x : 0.0; g[1] += s == 0 ? v.y : 0.0; g[2] += s == 0 ? v.z : 0.0;
    g[3] += s == 1 ? v.x : 0.0; g[4] += s == 1 ? v.y : 0.0; g[5] += s == 1 ? v.z : 0.0;
    g[6] += s == 2 ? v.x : 0.0; g[7] += s == 2 ? v.y : 0.0; g[8] += s == 2 ? v.z : 0.0;
    g[9] += s == 3 ? v.x : 0.0; g[10] += s == 3 ? v.y : 0.0; g[11] += s == 3 ? v.z : 0.0;
}

// squared distances with gradients (slots in the decoded vertex order)
__device__ inline double d_pp(V3 a, V3 b, double* g) {
    V3 u = a - b;
    if (g) { put(g, 0, 2.0 * u); put(g, 1, -2.0 * u); }
    return dot(u, u);
}
__device__ inline double d_pe(V3 p, V3 a, V3 b, double* g) {
    V3 u = a - p, v = b - p, n = cross(u, v), e = v - u;
    double nn = dot(n, n), L = dot(e, e), d = nn / L;
    if (g) {
        V3 du = (2.0 / L) * cross(v, n) + (2.0 * d / L) * e;
        V3 dv = (2.0 / L) * cross(n, u) - (2.0 * d / L) * e;
        put(g, 1, du); put(g, 2, dv); put(g, 0, -1.0 * (du + dv));
    }
    return d;
}
__device__ inline double d_pt(V3 p, V3 a, V3 b, V3 c, double* g) {
    V3 w = p - a, e1 = b - a, e2 = c - a, n = cross(e1, e2);
    double nn = dot(n, n), t = dot(w, n), d = t * t / nn;
    if (g) {
        V3 dw = (2.0 * t / nn) * n;
        V3 de1 = (2.0 * t / nn) * cross(e2, w) - (2.0 * d / nn) * cross(e2, n);
        V3 de2 = (2.0 * t / nn) * cross(w, e1) - (2.0 * d / nn) * cross(n, e1);
        put(g, 0, dw); put(g, 2, de1); put(g, 3, de2); put(g, 1, -1.0 * (dw + de1 + de2));
    }
    return d;
}
__device__ inline double d_ee(V3 a, V3 b, V3 c, V3 dd, double* g) {
    V3 w = c - a, e0 = b - a, e1 = dd - c, n = cross(e0, e1);
    double nn = dot(n, n), t = dot(w, n), d = t * t / nn;
    if (g) {
        V3 dw = (2.0 * t / nn) * n;
        V3 de0 = (2.0 * t / nn) * cross(e1, w) - (2.0 * d / nn) * cross(e1, n);
        V3 de1 = (2.0 * t / nn) * cross(w, e0) - (2.0 * d / nn) * cross(n, e0);
        put(g, 0, -1.0 * (dw + de0)); put(g, 1, de0); put(g, 2, dw - de1); put(g, 3, de1);
    }
    return d;
}

__device__ inline double bfun(double I, double k2) { double l = log(I); return k2 * (I - 1) * (I - 1) * l * l; }
__device__ inline double pk1(double I, double k2) {
    double l = log(I);
    return 2.0 * (2.0 * k2 * l * (I - 1) * (I + I * l - 1)) / I;
}
__device__ inline double lam0(double I, double k2) {
    double l = log(I);
    return -(4.0 * k2 * (4 * I + l - 3 * I * I * l * l + 6 * I * l - 2 * I * I + I * l * l - 7 * I * I * l - 2)) / I;
}

// decoded pair: vertices, count, size class, kinds
struct Pair {
    int v[4];
    int nv, cls, dkind, moll;
    int sl[4];  // mollifier edge slots (eA0, eA1, eB0, eB1)
};

__device__ inline Pair decode(int4 q) {
    Pair P;
    P.moll = 0;
    P.sl[0] = P.sl[1] = P.sl[2] = P.sl[3] = 0;
    int a = q.x, b = q.y, c = q.z, w = q.w;
    if (a >= 0) {
        P.v[0] = a; P.v[1] = b; P.v[2] = c; P.nv = 4; P.dkind = 3; P.cls = 4;
        if (w >= 0) P.v[3] = w;
        else { P.v[3] = -w - 1; P.moll = 1; P.sl[0] = 0; P.sl[1] = 1; P.sl[2] = 2; P.sl[3] = 3; }
    } else {
        int v0 = -a - 1;
        if (c < 0) {
            if (b < 0) {
                P.v[0] = v0; P.v[1] = -b - 1; P.v[2] = -c - 1; P.v[3] = -w - 1; P.nv = 4; P.dkind = 0; P.cls = 4;
                P.moll = 1; P.sl[0] = 0; P.sl[1] = 2; P.sl[2] = 1; P.sl[3] = 3;
            } else { P.v[0] = v0; P.v[1] = b; P.v[2] = P.v[3] = 0; P.nv = 2; P.dkind = 0; P.cls = 2; }
        } else if (w < 0) {
            if (b < 0) {
                P.v[0] = v0; P.v[1] = -b - 1; P.v[2] = c; P.v[3] = -w - 1; P.nv = 4; P.dkind = 1; P.cls = 4;
                P.moll = 1; P.sl[0] = 0; P.sl[1] = 3; P.sl[2] = 1; P.sl[3] = 2;
            } else { P.v[0] = v0; P.v[1] = b; P.v[2] = c; P.v[3] = 0; P.nv = 3; P.dkind = 1; P.cls = 3; }
        } else { P.v[0] = v0; P.v[1] = b; P.v[2] = c; P.v[3] = w; P.nv = 4; P.dkind = 2; P.cls = 4; }
    }
    return P;
}

// The per-pair Hessian form: entry (i, j) = gc_i (m00 gc_j + m01 gs_j) +
// gs_i (m01 gc_j + m11 gs_j), evaluated where the blocks are written, so the
// 12x12 matrix never lives in registers.
struct HForm {
    double gs[12], gc[12];
    double m00, m01, m11;  // moll: 2x2 PSD; non-moll: m11 = lambda0, others 0
};

constexpr int kBarT = 128;  // pairs (threads) per k_barrier workgroup

__global__ void __launch_bounds__(kBarT) k_barrier(int np, const int4* __restrict__ pairs, const int* __restrict__ mat_index,
                                                 int c2, int c3, int c4, const double* __restrict__ x,
                                                 const double* __restrict__ xr, double dhat, double kappa,
                                                 const int* __restrict__ btype, int index_offset,
                                                 double* __restrict__ fb, double* __restrict__ tv,
                                                 int* __restrict__ tr, int* __restrict__ tc,
                                                 double* __restrict__ gscr, unsigned* __restrict__ gkey,
                                                 unsigned* __restrict__ grec, unsigned gsentinel) {
    // one thread per pair computes its barrier; the workgroup then writes the
    // pairs' triplet blocks together (below)
    __shared__ double sF[27][kBarT];  // per pair: gs[12], gc[12], m00, m01, m11
    __shared__ int sV[4][kBarT];      // vertices
    __shared__ int sInfo[kBarT];      // nv | pinned-vertex mask << 4 | active << 8; 0: no pair
    __shared__ long long sOff[kBarT];  // first triplet of the pair
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < np;
    if (valid) {
        Pair P = decode(pairs[p]);
        V3 X[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) X[a] = (a < P.nv) ? ld(x, P.v[a]) : V3{0, 0, 0};
        double gd[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) gd[k] = 0.0;
        double d;
        if (P.dkind == 0) d = d_pp(X[0], X[1], gd);
        else if (P.dkind == 1) d = d_pe(X[0], X[1], X[2], gd);
        else if (P.dkind == 2) d = d_pt(X[0], X[1], X[2], X[3], gd);
        else d = d_ee(X[0], X[1], X[2], X[3], gd);
        const double k2 = kappa * dhat * dhat;
        double I = d / dhat, s = sqrt(I);
        double inv = (d > 0.0) ? 1.0 / (2.0 * sqrt(d * dhat)) : 0.0;
        HForm F;
#pragma unroll
        for (int k = 0; k < 12; ++k) { F.gs[k] = gd[k] * inv; F.gc[k] = 0.0; }
        double grad[12];
        bool active = true;
        if (!P.moll) {
            double pp = pk1(I, k2) * s;
#pragma unroll
            for (int k = 0; k < 12; ++k) grad[k] = F.gs[k] * pp;
            F.m00 = 0.0; F.m01 = 0.0;
            F.m11 = lam0((d < 1e-6 * dhat) ? 1e-6 : I, k2);
        } else {
            V3 eA = pick4(X, P.sl[1]) - pick4(X, P.sl[0]), eB = pick4(X, P.sl[3]) - pick4(X, P.sl[2]);
            V3 v = cross(eA, eB);
            double cc = sqrt(dot(v, v)), I1 = cc * cc;
            if (I1 == 0.0) {
                active = false;
#pragma unroll
                for (int k = 0; k < 12; ++k) grad[k] = 0.0;
                F.m00 = F.m01 = F.m11 = 0.0;
            } else {
                V3 vh = (1.0 / cc) * v;
                V3 dA = cross(eB, vh), dB = cross(vh, eA);
                add_at(F.gc, P.sl[1], dA); add_at(F.gc, P.sl[0], -1.0 * dA);
                add_at(F.gc, P.sl[3], dB); add_at(F.gc, P.sl[2], -1.0 * dB);
                V3 r0 = ld(xr, pick4(P.v, P.sl[0])) - ld(xr, pick4(P.v, P.sl[1]));
                V3 r1 = ld(xr, pick4(P.v, P.sl[2])) - ld(xr, pick4(P.v, P.sl[3]));
                double eps = 1e-3 * dot(r0, r0) * dot(r1, r1);
                double l = log(I), e2 = eps * eps;
                double p1 = -kappa * 2.0 * (2.0 * dhat * dhat * l * l * (I1 - eps) * (I - 1) * (I - 1)) / e2;
                double p2 = -kappa * 2.0 * (2.0 * I1 * dhat * dhat * l * (I1 - 2 * eps) * (I - 1) * (I + I * l - 1)) / (I * e2);
#pragma unroll
                for (int k = 0; k < 12; ++k) grad[k] = F.gc[k] * cc * p1 + F.gs[k] * s * p2;
                double l10 = -kappa * (4.0 * dhat * dhat * l * l * (I - 1) * (I - 1) * (3 * I1 - eps)) / e2;
                double l20 = kappa *
                             (4.0 * I1 * dhat * dhat * (I1 - 2 * eps) *
                              (4 * I + l - 3 * I * I * l * l + 6 * I * l - 2 * I * I + I * l * l - 7 * I * I * l - 2)) /
                             (I * e2);
                double lg = -kappa * 4.0 * cc * s * (4.0 * dhat * dhat * l * (I1 - eps) * (I - 1) * (I + I * l - 1)) / (I * e2);
                double m00 = l10, m01 = lg, m11 = l20;
                double trc = m00 + m11, det = m00 * m11 - m01 * m01;
                double disc = sqrt(fmax(0.25 * trc * trc - det, 0.0));
                double ev0 = 0.5 * trc - disc, ev1 = 0.5 * trc + disc;
                if (ev0 < 0.0) {
                    double vx, vy;
                    if (fabs(m01) > 0.0) { vx = ev0 - m11; vy = m01; }
                    else if (m00 <= m11) { vx = 1.0; vy = 0.0; }
                    else { vx = 0.0; vy = 1.0; }
                    double nrm = sqrt(vx * vx + vy * vy);
                    vx /= nrm; vy /= nrm;
                    double ux = -vy, uy = vx, pe = ev1 > 0.0 ? ev1 : 0.0;
                    m00 = pe * ux * ux; m01 = pe * ux * uy; m11 = pe * uy * uy;
                }
                F.m00 = m00; F.m01 = m01; F.m11 = m11;
            }
        }
        // gradient: FP64 atomics, as the reference; or, in the ordered mode, one
        // record per pair vertex for a sorted per-vertex sum (k_grad_gather)
        if (gscr) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const bool use = active && a < P.nv;
#pragma unroll
                for (int k = 0; k < 3; ++k) gscr[12 * (size_t)p + 3 * a + k] = use ? grad[3 * a + k] : 0.0;
                gkey[4 * (size_t)p + a] = use ? (unsigned)P.v[a] : gsentinel;
                grec[4 * (size_t)p + a] = 4u * (unsigned)p + (unsigned)a;
            }
        } else if (active) {
#pragma unroll
            for (int a = 0; a < 4; ++a)
                if (a < P.nv)
#pragma unroll
                    for (int k = 0; k < 3; ++k) atomicAdd(fb + 3 * (size_t)P.v[a] + k, grad[3 * a + k]);
        }
        // triplet block layout: the pair's nv x nv blocks at off
        long long off;
        int mi = mat_index[p];
        if (P.cls == 4) off = 16LL * mi;
        else if (P.cls == 3) off = 16LL * c4 + 9LL * mi;
        else off = 16LL * c4 + 9LL * c3 + 4LL * mi;
        int pinmask = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
            if (a < P.nv && btype && btype[P.v[a]] != 0) pinmask |= 1 << a;
        const int t = threadIdx.x;
#pragma unroll
        for (int k = 0; k < 12; ++k) { sF[k][t] = F.gs[k]; sF[12 + k][t] = F.gc[k]; }
        sF[24][t] = F.m00; sF[25][t] = F.m01; sF[26][t] = F.m11;
#pragma unroll
        for (int a = 0; a < 4; ++a) sV[a][t] = P.v[a];
        sInfo[t] = P.nv | (pinmask << 4) | ((active ? 1 : 0) << 8);
        sOff[t] = off;
    } else {
        sInfo[threadIdx.x] = 0;
    }
    __syncthreads();
    // Triplets, written by the workgroup together: work item w = (pair slot
    // w / 16, block w % 16), so consecutive lanes write consecutive 72-byte
    // blocks of one pair.  Block kk = ii nv + jj of the nv x nv grid holds
    // (v_ii, v_jj); blocks below the diagonal, of a pinned vertex or of an
    // inactive pair are written as zeros (then the upper/boundary remap of
    // the indices).
    for (int w = threadIdx.x; w < 16 * kBarT; w += kBarT) {
        const int slot = w >> 4, kk = w & 15;
        const int info = sInfo[slot];
        const int nv = info & 15;
        if (kk >= nv * nv) continue;
        const int ii = kk / nv, jj = kk - ii * nv;
        const int r = sV[ii][slot], c = sV[jj][slot];
        const bool keep = r <= c;
        const bool zero = !keep || ((info >> 4) & ((1 << ii) | (1 << jj))) || !((info >> 8) & 1);
        const long long o = sOff[slot] + kk;
        double* B = tv + 9 * o;
#pragma unroll
        for (int qq = 0; qq < 3; ++qq)
#pragma unroll
            for (int pp = 0; pp < 3; ++pp) {
                double h = 0.0;
                if (!zero) {
                    const int i = 3 * ii + pp, j = 3 * jj + qq;
                    const double gci = sF[12 + i][slot], gsi = sF[i][slot];
                    const double gcj = sF[12 + j][slot], gsj = sF[j][slot];
                    const double m00 = sF[24][slot], m01 = sF[25][slot], m11 = sF[26][slot];
                    h = gci * (m00 * gcj + m01 * gsj) + gsi * (m01 * gcj + m11 * gsj);
                }
                B[pp + 3 * qq] = h;
            }
        tr[o] = (keep ? r : c) + index_offset;
        tc[o] = (keep ? c : r) + index_offset;
    }
}

// ---- narrow phase: classification + encoding ---------------------------
// Energy of one pair, RANK 2: kappa dhat^2 (I-1)^2 ln^2 I, times the
// mollifier -I1 (I1 - 2 eps) / eps^2 for nearly parallel edges
// (__cal_Barrier_energy, S/GIPC.cu:701-957); same distances as k_barrier.
__device__ double pair_energy(int4 q, const double* __restrict__ x, const double* __restrict__ xr, double dhat,
                              double kappa) {
    Pair P = decode(q);
    V3 X[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) X[a] = (a < P.nv) ? ld(x, P.v[a]) : V3{0, 0, 0};
    double d;
    if (P.dkind == 0) d = d_pp(X[0], X[1], nullptr);
    else if (P.dkind == 1) d = d_pe(X[0], X[1], X[2], nullptr);
    else if (P.dkind == 2) d = d_pt(X[0], X[1], X[2], X[3], nullptr);
    else d = d_ee(X[0], X[1], X[2], X[3], nullptr);
    const double I = d / dhat, b = bfun(I, kappa * dhat * dhat);
    if (!P.moll) return b;
    V3 v = cross(pick4(X, P.sl[1]) - pick4(X, P.sl[0]), pick4(X, P.sl[3]) - pick4(X, P.sl[2]));
    double cc = sqrt(dot(v, v)), I1 = cc * cc;
    if (I1 == 0.0) return 0.0;  // no contribution, as in the reference
    V3 r0 = ld(xr, pick4(P.v, P.sl[0])) - ld(xr, pick4(P.v, P.sl[1]));
    V3 r1 = ld(xr, pick4(P.v, P.sl[2])) - ld(xr, pick4(P.v, P.sl[3]));
    double eps = 1e-3 * dot(r0, r0) * dot(r1, r1);
    return -I1 * (I1 - 2 * eps) / (eps * eps) * b;
}

__global__ void __launch_bounds__(256) k_barrier_energy(int np, const int4* __restrict__ pairs,
                                                        const double* __restrict__ x, const double* __restrict__ xr,
                                                        double dhat, double kappa, double* __restrict__ part) {
    __shared__ double sh[4];
    double e = 0.0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x)
        e += pair_energy(pairs[p], x, xr, dhat, kappa);
    write_partial(e, part, sh);
}

__device__ inline double det3c(V3 c0, V3 c1, V3 c2) { return dot(c0, cross(c1, c2)); }
__device__ inline void pe_param(V3 v0, V3 a, V3 b, V3 nvec, double& px, double& py) {
    V3 b0 = b - a, b1 = cross(b0, nvec), b2 = v0 - a;
    double D = det3c(b0, b1, nvec);
    px = det3c(b2, b1, nvec) / D;
    py = det3c(b0, b2, nvec) / D;
}
__device__ inline int dtype_pt(V3 v0, V3 v1, V3 v2, V3 v3) {
    V3 nvec = cross(v2 - v1, v3 - v1);
    double p0x, p0y, p1x, p1y, p2x, p2y;
    pe_param(v0, v1, v2, nvec, p0x, p0y);
    if (p0x > 0 && p0x < 1 && p0y >= 0) return 3;
    pe_param(v0, v2, v3, nvec, p1x, p1y);
    if (p1x > 0.0 && p1x < 1.0 && p1y >= 0.0) return 4;
    pe_param(v0, v3, v1, nvec, p2x, p2y);
    if (p2x > 0.0 && p2x < 1.0 && p2y >= 0.0) return 5;
    if (p0x <= 0.0 && p2x >= 1.0) return 0;
    if (p1x <= 0.0 && p0x >= 1.0) return 1;
    if (p2x <= 0.0 && p1x >= 1.0) return 2;
    return 6;
}
__device__ inline int dtype_ee(V3 v0, V3 v1, V3 v2, V3 v3) {
    V3 u = v1 - v0, v = v3 - v2, w = v0 - v2;
    double a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
    double D = a * c - b * b, tD = D, sN, tN;
    int def = 8;
    sN = (b * e - c * d);
    if (sN <= 0.0) { tN = e; tD = c; def = 2; }
    else if (sN >= D) { tN = e + b; tD = c; def = 5; }
    else {
        tN = (a * e - b * d);
        V3 uv = cross(u, v);
        if (tN > 0.0 && tN < tD && (dot(w, uv) == 0.0 || dot(uv, uv) < 1.0e-20 * a * c)) {
            if (sN < D / 2) { tN = e; tD = c; def = 2; }
            else { tN = e + b; tD = c; def = 5; }
        }
    }
    if (tN <= 0.0) {
        if (-d <= 0.0) return 0;
        else if (-d >= a) return 3;
        else return 6;
    } else if (tN >= tD) {
        if ((-d + b) <= 0.0) return 1;
        else if ((-d + b) >= a) return 4;
        else return 7;
    }
    return def;
}

// classify candidate i (PT candidates first, then EE): encoded pair + class
// (2,3,4) or class 0 when inactive
__global__ void k_classify(int npt, const int4* __restrict__ pt, int nee, const int4* __restrict__ ee,
                           const double* __restrict__ x, const double* __restrict__ xr, double dhat,
                           int4* __restrict__ enc, int* __restrict__ cls) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npt + nee) return;
    int4 out = make_int4(0, 0, 0, 0);
    int k = 0;
    if (i < npt) {
        int4 c = pt[i];
        V3 v0 = ld(x, c.x), v1 = ld(x, c.y), v2 = ld(x, c.z), v3 = ld(x, c.w);
        int t = dtype_pt(v0, v1, v2, v3);
        double d;
        switch (t) {
            case 0: d = d_pp(v0, v1, nullptr); if (d < dhat) { out = make_int4(-c.x - 1, c.y, -1, -1); k = 2; } break;
            case 1: d = d_pp(v0, v2, nullptr); if (d < dhat) { out = make_int4(-c.x - 1, c.z, -1, -1); k = 2; } break;
            case 2: d = d_pp(v0, v3, nullptr); if (d < dhat) { out = make_int4(-c.x - 1, c.w, -1, -1); k = 2; } break;
            case 3: d = d_pe(v0, v1, v2, nullptr); if (d < dhat) { out = make_int4(-c.x - 1, c.y, c.z, -1); k = 3; } break;
            case 4: d = d_pe(v0, v2, v3, nullptr); if (d < dhat) { out = make_int4(-c.x - 1, c.z, c.w, -1); k = 3; } break;
            case 5: d = d_pe(v0, v3, v1, nullptr); if (d < dhat) { out = make_int4(-c.x - 1, c.w, c.y, -1); k = 3; } break;
            default: d = d_pt(v0, v1, v2, v3, nullptr); if (d < dhat) { out = make_int4(-c.x - 1, c.y, c.z, c.w); k = 4; } break;
        }
    } else {
        int j = i - npt;
        int4 c = ee[j];
        V3 v0 = ld(x, c.x), v1 = ld(x, c.y), v2 = ld(x, c.z), v3 = ld(x, c.w);
        int t = dtype_ee(v0, v1, v2, v3);
        V3 r0 = ld(xr, c.x) - ld(xr, c.y), r1 = ld(xr, c.z) - ld(xr, c.w);
        double eps = 1e-3 * dot(r0, r0) * dot(r1, r1);
        V3 cr = cross(v0 - v1, v2 - v3);
        int ae = (dot(cr, cr) < eps) ? -j - 2 : -1;
        double d;
        switch (t) {
            case 0: d = d_pp(v0, v2, nullptr); if (d < dhat) { out = make_int4(-c.x - 1, c.z, -1, ae); k = 2; } break;
            case 1: d = d_pp(v0, v3, nullptr); if (d < dhat) { out = make_int4(-c.x - 1, c.w, -1, ae); k = 2; } break;
            case 2: d = d_pe(v0, v2, v3, nullptr); if (d < dhat) { out = make_int4(-c.x - 1, c.z, c.w, ae); k = 3; } break;
            case 3: d = d_pp(v1, v2, nullptr); if (d < dhat) { out = make_int4(-c.y - 1, c.z, -1, ae); k = 2; } break;
            case 4: d = d_pp(v1, v3, nullptr); if (d < dhat) { out = make_int4(-c.y - 1, c.w, -1, ae); k = 2; } break;
            case 5: d = d_pe(v1, v2, v3, nullptr); if (d < dhat) { out = make_int4(-c.y - 1, c.z, c.w, ae); k = 3; } break;
            case 6: d = d_pe(v2, v0, v1, nullptr); if (d < dhat) { out = make_int4(-c.z - 1, c.x, c.y, ae); k = 3; } break;
            case 7: d = d_pe(v3, v0, v1, nullptr); if (d < dhat) { out = make_int4(-c.w - 1, c.x, c.y, ae); k = 3; } break;
            default: d = d_ee(v0, v1, v2, v3, nullptr); if (d < dhat) { out = make_int4(c.x, c.y, c.z, c.w); k = 4; } break;
        }
    }
    enc[i] = out;
    cls[i] = k;
}

// flags[c][i] = (cls[i] == c) for c = 2,3,4 and active; laid out as 4 rows
__global__ void k_class_flags(int m, const int* __restrict__ cls, int* __restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int k = cls[i];
    flags[i] = k != 0;                       // row 0: any active
    flags[(size_t)(m + 1) + i] = k == 2;     // row 1
    flags[2 * (size_t)(m + 1) + i] = k == 3; // row 2
    flags[3 * (size_t)(m + 1) + i] = k == 4; // row 3
}

__global__ void k_class_scatter(int m, const int4* __restrict__ enc, const int* __restrict__ cls,
                                const int* __restrict__ pos, int4* __restrict__ pairs, int* __restrict__ mat_index) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int k = cls[i];
    if (k == 0) return;
    int o = pos[i];
    pairs[o] = enc[i];
    mat_index[o] = pos[(size_t)(k - 1) * (m + 1) + i];
}

// ---- additive continuous collision detection ---------------------------------
// Squared distance of the closest feature pair (the classification above).
__device__ inline double dist2_pt_any(V3 p, V3 a, V3 b, V3 c) {
    switch (dtype_pt(p, a, b, c)) {
        case 0: return d_pp(p, a, nullptr);
        case 1: return d_pp(p, b, nullptr);
        case 2: return d_pp(p, c, nullptr);
        case 3: return d_pe(p, a, b, nullptr);
        case 4: return d_pe(p, b, c, nullptr);
        case 5: return d_pe(p, c, a, nullptr);
        default: return d_pt(p, a, b, c, nullptr);
    }
}
__device__ inline double dist2_ee_any(V3 a0, V3 a1, V3 b0, V3 b1) {
    switch (dtype_ee(a0, a1, b0, b1)) {
        case 0: return d_pp(a0, b0, nullptr);
        case 1: return d_pp(a0, b1, nullptr);
        case 2: return d_pe(a0, b0, b1, nullptr);
        case 3: return d_pp(a1, b0, nullptr);
        case 4: return d_pp(a1, b1, nullptr);
        case 5: return d_pe(a1, b0, b1, nullptr);
        case 6: return d_pe(b0, a0, a1, nullptr);
        case 7: return d_pe(b1, a0, a1, nullptr);
        default: return d_ee(a0, a1, b0, b1, nullptr);
    }
}
__device__ inline double dist2_ee_safe(V3 a0, V3 a1, V3 b0, V3 b1) {
    double d = dist2_ee_any(a0, a1, b0, b1);
    if (d <= 0.0)  // touching or degenerate: the closest endpoint pair
        d = fmin(fmin(d_pp(a0, b0, nullptr), d_pp(a0, b1, nullptr)), fmin(d_pp(a1, b0, nullptr), d_pp(a1, b1, nullptr)));
    return d;
}
constexpr int kAccdMaxSteps = 50000;

// Additive CCD (Li et al. 2021, "Codimensional incremental potential
// contact"): advance by conservative steps (1 - eta) d / (max relative
// displacement) until the distance falls below eta of its start; time of
// impact as a fraction of the step, capped at 1.  Four vertices v[0..3]
// moving by u[0..3]; pt selects point-triangle (else edge-edge).
// Reference: ACCD.cu point_triangle_ccd :456-534, edge_edge_ccd :373-454,
// thickness 0 as GIPC::self_largestFeasibleStepSize uses them.
__device__ double accd(bool pt, V3 v[4], V3 u[4], double eta) {
    const V3 mov = -0.25 * (u[0] + u[1] + u[2] + u[3]);
#pragma unroll
    for (int i = 0; i < 4; ++i) u[i] = u[i] + mov;
    double maxd;
    if (pt) {
        maxd = sqrt(dot(u[0], u[0])) + sqrt(fmax(dot(u[1], u[1]), fmax(dot(u[2], u[2]), dot(u[3], u[3]))));
    } else {
        maxd = sqrt(fmax(dot(u[0], u[0]), dot(u[1], u[1]))) + sqrt(fmax(dot(u[2], u[2]), dot(u[3], u[3])));
    }
    if (maxd == 0.0) return 1.0;
    auto d2 = [&](void) {
        return pt ? dist2_pt_any(v[0], v[1], v[2], v[3]) : dist2_ee_safe(v[0], v[1], v[2], v[3]);
    };
    double dd = d2();
    double dist = sqrt(dd);
    const double gap = eta * dd / dist;
    double toc = 0.0;
    for (int it = 0; it < kAccdMaxSteps; ++it) {
        const double lb = (1.0 - eta) * dd / (dist * maxd);
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = v[i] + lb * u[i];
        dd = d2();
        dist = sqrt(dd);
        if (toc != 0.0 && dd / dist < gap) return toc;
        toc += lb;
        if (toc > 1.0) return 1.0;
    }
    return toc;
}

// per candidate time of impact along -dir; block minima (1 if none)
__global__ void __launch_bounds__(256) k_ccd(int npt, const int4* __restrict__ pt, int nee, const int4* __restrict__ ee,
                                             const double* __restrict__ x, const double* __restrict__ dir, double eta,
                                             double* __restrict__ part) {
    __shared__ double sh[256 / 64];
    double t = 1.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < npt + nee; i += gridDim.x * blockDim.x) {
        const bool isp = i < npt;
        const int4 q = isp ? pt[i] : ee[i - npt];
        const int id[4] = {q.x, q.y, q.z, q.w};
        V3 v[4], u[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = ld(x, id[k]);
            u[k] = -1.0 * ld(dir, id[k]);
        }
        t = fmin(t, accd(isp, v, u, eta));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t = fmin(t, __shfl_xor(t, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sh[0];
        for (int w = 1; w < 4; ++w) m = fmin(m, sh[w]);
        part[blockIdx.x] = m;
    }
}

__global__ void __launch_bounds__(256) k_min_total(const double* __restrict__ part, int np, double* __restrict__ out) {
    __shared__ double sh[256 / 64];
    double t = 1.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) t = fmin(t, part[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t = fmin(t, __shfl_xor(t, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sh[0];
        for (int w = 1; w < 4; ++w) m = fmin(m, sh[w]);
        *out = m;
    }
}

// fb[v] += the sorted records of vertex v, summed in pair order (one head
// thread per vertex)
__global__ void k_grad_gather(long long m, const unsigned* __restrict__ key, const unsigned* __restrict__ rec,
                              const double* __restrict__ gscr, unsigned sentinel, double* __restrict__ fb) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const unsigned v = key[i];
    if (v == sentinel || (i > 0 && key[i - 1] == v)) return;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (long long j = i; j < m && key[j] == v; ++j) {
        const double* g = gscr + 3 * (size_t)rec[j];  // record 4 p + a -> gradient row 3 (4 p + a)
        g0 += g[0]; g1 += g[1]; g2 += g[2];
    }
    fb[3 * (size_t)v] += g0;
    fb[3 * (size_t)v + 1] += g1;
    fb[3 * (size_t)v + 2] += g2;
}

}  // namespace

void barrier_assemble(Ctx& c, int np, const int* pairs, const int* mat_index, const int counts[3], const double* x,
                      const double* xrest, double dhat, double kappa, const int* btype, int index_offset,
                      long long triplet_offset, double* fb) {
    if (!c.t_vals) throw state_error("barrier: no triplet view set");
    long long need = 16LL * counts[2] + 9LL * counts[1] + 4LL * counts[0];
    if (triplet_offset < 0 || triplet_offset + need > c.t_cap) throw capacity_error("barrier: triplet capacity exceeded");
    if (np <= 0) return;
    prof_begin(c, PROF_BARRIER);
    double* gscr = nullptr;
    unsigned *gkey = nullptr, *grec = nullptr;
    const unsigned sentinel = 0xFFFFFFFFu;  // sorts after every vertex id
    const long long m = 4LL * np;
    if (c.deterministic) {
        gscr = c.grad_scr.reserve(12 * (size_t)np);
        gkey = c.grad_key.reserve(2 * (size_t)m);
        grec = c.grad_rec.reserve(2 * (size_t)m);
    }
    hipLaunchKernelGGL(k_barrier, dim3(div_up(np, 128)), dim3(128), 0, c.stream, np, (const int4*)pairs, mat_index,
                       counts[0], counts[1], counts[2], x, xrest, dhat, kappa, btype, index_offset, fb,
                       c.t_vals + 9 * triplet_offset, c.t_rows + triplet_offset, c.t_cols + triplet_offset, gscr,
                       gkey, grec, sentinel);
    SGIPC_LAUNCH_CHECK();
    if (c.deterministic) {
        if (c.sort_impl == SORT_OWN) {
            radix_sort_pairs(c, gkey, gkey + m, (const int*)grec, 0, (int*)(grec + m), m, 32, c.stream);
        } else {  // rocPRIM as the cross-check
            size_t bytes = 0;
            SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, gkey, gkey + m, grec, grec + m, (size_t)m, 0, 32,
                                                c.stream));
            void* tmp = c.grad_sort_tmp.reserve(bytes + 1);
            SGIPC_HIP(rocprim::radix_sort_pairs(tmp, bytes, gkey, gkey + m, grec, grec + m, (size_t)m, 0, 32,
                                                c.stream));
        }
        hipLaunchKernelGGL(k_grad_gather, dim3(div_up(m, 256)), dim3(256), 0, c.stream, m,
                           (const unsigned*)(gkey + m), (const unsigned*)(grec + m), (const double*)gscr, sentinel,
                           fb);
        SGIPC_LAUNCH_CHECK();
    }
    prof_end(c, PROF_BARRIER);
}

int classify_contacts(Ctx& c, int npt, const int* pt, int nee, const int* ee, const double* x, const double* xrest,
                      double dhat, int* pairs, int* mat_index, int capacity, int counts[3]) {
    hipStream_t s = c.stream;
    int m = npt + nee;
    counts[0] = counts[1] = counts[2] = 0;
    if (m <= 0) return 0;
    int4* enc = (int4*)c.cls_buf.reserve(4 * (size_t)m + 4);
    int* cls = c.cls_flag.reserve((size_t)m + 1);
    int* flags = c.cls_scan.reserve(4 * ((size_t)m + 1) + 4);
    int* pos = c.cls_pos.reserve(4 * ((size_t)m + 1) + 4);
    int* tot = c.dev_int.reserve(8);
    hipLaunchKernelGGL(k_classify, dim3(div_up(m, 256)), dim3(256), 0, s, npt, (const int4*)pt, nee, (const int4*)ee, x,
                       xrest, dhat, enc, cls);
    SGIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_class_flags, dim3(div_up(m, 256)), dim3(256), 0, s, m, cls, flags);
    SGIPC_LAUNCH_CHECK();
    for (int r = 0; r < 4; ++r) exclusive_scan(c, flags + (size_t)r * (m + 1), pos + (size_t)r * (m + 1), m, nullptr, s);
    int hcnt[4];
    for (int r = 0; r < 4; ++r)
        SGIPC_HIP(hipMemcpyAsync(&hcnt[r], pos + (size_t)r * (m + 1) + m, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    (void)tot;
    if (hcnt[0] > capacity) throw capacity_error("classify: pair capacity exceeded");
    hipLaunchKernelGGL(k_class_scatter, dim3(div_up(m, 256)), dim3(256), 0, s, m, enc, cls, pos, (int4*)pairs, mat_index);
    SGIPC_LAUNCH_CHECK();
    counts[0] = hcnt[1];
    counts[1] = hcnt[2];
    counts[2] = hcnt[3];
    return hcnt[0];
}

double ccd_step_bound(Ctx& c, int npt, const int* pt, int nee, const int* ee, const double* x, const double* dir,
                      double slack) {
    if (!(slack > 0.0 && slack < 1.0)) throw arg_error("ccd_step_bound: slackness must be in (0, 1)");
    const int m = npt + nee;
    if (m <= 0) return 1.0;
    double* part = energy_partials(c);
    const int g = energy_grid(m);
    hipLaunchKernelGGL(k_ccd, dim3(g), dim3(256), 0, c.stream, npt, (const int4*)pt, nee, (const int4*)ee, x, dir,
                       1.0 - slack, part);
    SGIPC_LAUNCH_CHECK();
    double* out = part + g;
    hipLaunchKernelGGL(k_min_total, dim3(1), dim3(256), 0, c.stream, (const double*)part, g, out);
    SGIPC_LAUNCH_CHECK();
    double h = 1.0;
    SGIPC_HIP(hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SGIPC_HIP(hipStreamSynchronize(c.stream));
    return h;
}

double barrier_energy(Ctx& c, int np, const int* pairs, const double* x, const double* xr, double dhat,
                      double kappa) {
    double* part = energy_partials(c);
    int g = energy_grid(np);
    hipLaunchKernelGGL(k_barrier_energy, dim3(g), dim3(256), 0, c.stream, np, (const int4*)pairs, x, xr, dhat, kappa,
                       part);
    SGIPC_LAUNCH_CHECK();
    return energy_finish(c, part, g);
}

}  // namespace sgipc
