/*
 * sgipc_oracle_barrier.c -- CPU restatement of the IPC barrier assembly and
 * of the narrow-phase pair classification.  TEST INFRASTRUCTURE ONLY (see
 * sgipc_oracle.c).
 *
 * Barrier (S/GIPC.cu:3106-4943, RANK 2, NEWF): with d the squared distance of
 * the pair's sub-case and I = d / dhat,
 *     b(I) = kappa dhat^2 (I - 1)^2 ln^2 I                  (:701-957)
 * Non-mollified PT / EE / PE / PP: s = sqrt(I), g = ds/dx,
 *     grad = g * s * 2 b'(I)        (flatten_pk1, :3183-3186)
 *     H    = lambda0 g g^T, lambda0 = d^2 b / ds^2 evaluated at
 *            max(I, 1e-6) when d < 1e-6 dhat            (:3264-3284)
 * The reference obtains g from generated code whose only non-zero column is
 * ds/dx (S/GIPC_PDerivative.cuh:1861-2406); here g is the analytic gradient
 * of the distance formulas of S/mlbvh.cu:229-257.
 * Mollified EE / PE / PP (pairs with a negative marker): e(I1) = I1 (2 eps -
 * I1) / eps^2 with I1 = c^2, c = |eA x eB|, eps = 1e-3 |eA0|^2 |eB0|^2 (rest
 * edges, S/mlbvh.cu:273-277); energy e(I1) b(I2); gradient and a Gauss-Newton
 * Hessian in the two scalars (c, s): [gc gs] PSD([[E_cc, E_cs],[E_cs, E_ss]])
 * [gc gs]^T -- the generated derivatives of :1616-1860 have non-zero
 * columns only for c (column 4) and s (column 8), so the lambda11/lambda12
 * terms of :3481-3500 vanish.
 */
#include "sgipc_oracle.h"
#include <math.h>
#include <string.h>

static void sub3(const double* a, const double* b, double* o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static const double* X(const double* x, int i) { return x + 3 * (size_t)i; }

/* squared distances and their gradients w.r.t. the participating vertices */
static double d_pp(const double* a, const double* b, double* g /*6*/) {
    double u[3];
    sub3(a, b, u);
    if (g) for (int k = 0; k < 3; ++k) { g[k] = 2 * u[k]; g[3 + k] = -2 * u[k]; }
    return dot3(u, u);
}
/* point p, edge (a, b) */
static double d_pe(const double* p, const double* a, const double* b, double* g /*9: p,a,b*/) {
    double u[3], v[3], n[3], e[3];
    sub3(a, p, u); sub3(b, p, v); cross3(u, v, n); sub3(v, u, e);
    double nn = dot3(n, n), L = dot3(e, e), d = nn / L;
    if (g) {
        double vn[3], nu[3];
        cross3(v, n, vn); cross3(n, u, nu);
        for (int k = 0; k < 3; ++k) {
            double du = 2 * vn[k] / L + d * 2 * e[k] / L;  /* d/du */
            double dv = 2 * nu[k] / L - d * 2 * e[k] / L;  /* d/dv */
            g[3 + k] = du; g[6 + k] = dv; g[k] = -(du + dv);
        }
    }
    return d;
}
/* point p, triangle (a, b, c) */
static double d_pt(const double* p, const double* a, const double* b, const double* c, double* g /*12*/) {
    double w[3], e1[3], e2[3], n[3];
    sub3(p, a, w); sub3(b, a, e1); sub3(c, a, e2); cross3(e1, e2, n);
    double nn = dot3(n, n), t = dot3(w, n), d = t * t / nn;
    if (g) {
        double e2w[3], we1[3], e2n[3], ne1[3];
        cross3(e2, w, e2w); cross3(w, e1, we1); cross3(e2, n, e2n); cross3(n, e1, ne1);
        for (int k = 0; k < 3; ++k) {
            double dw = 2 * t * n[k] / nn;
            double de1 = 2 * t * e2w[k] / nn - 2 * d * e2n[k] / nn;
            double de2 = 2 * t * we1[k] / nn - 2 * d * ne1[k] / nn;
            g[k] = dw; g[6 + k] = de1; g[9 + k] = de2; g[3 + k] = -(dw + de1 + de2);
        }
    }
    return d;
}
/* edge (a, b), edge (c, d) */
static double d_ee(const double* a, const double* b, const double* c, const double* dd, double* g /*12*/) {
    double w[3], e0[3], e1[3], n[3];
    sub3(c, a, w); sub3(b, a, e0); sub3(dd, c, e1); cross3(e0, e1, n);
    double nn = dot3(n, n), t = dot3(w, n), d = t * t / nn;
    if (g) {
        double e1w[3], we0[3], e1n[3], ne0[3];
        cross3(e1, w, e1w); cross3(w, e0, we0); cross3(e1, n, e1n); cross3(n, e0, ne0);
        for (int k = 0; k < 3; ++k) {
            double dw = 2 * t * n[k] / nn;
            double de0 = 2 * t * e1w[k] / nn - 2 * d * e1n[k] / nn;
            double de1 = 2 * t * we0[k] / nn - 2 * d * ne0[k] / nn;
            g[k] = -dw - de0; g[3 + k] = de0; g[6 + k] = dw - de1; g[9 + k] = de1;
        }
    }
    return d;
}

/* c = |eA x eB| with eA = x[i1]-x[i0], eB = x[j1]-x[j0]; gradient into the
 * slots (si0, si1, sj0, sj1) of a 12-vector */
static double cross_norm(const double* xi0, const double* xi1, const double* xj0, const double* xj1, int si0, int si1,
                         int sj0, int sj1, double* g) {
    double eA[3], eB[3], v[3];
    sub3(xi1, xi0, eA); sub3(xj1, xj0, eB); cross3(eA, eB, v);
    double c = sqrt(dot3(v, v));
    if (g && c > 0) {
        double vh[3] = {v[0] / c, v[1] / c, v[2] / c}, dA[3], dB[3];
        cross3(eB, vh, dA); /* d|v|/d eA */
        cross3(vh, eA, dB); /* d|v|/d eB */
        for (int k = 0; k < 3; ++k) {
            g[3 * si1 + k] += dA[k]; g[3 * si0 + k] -= dA[k];
            g[3 * sj1 + k] += dB[k]; g[3 * sj0 + k] -= dB[k];
        }
    }
    return c;
}

static double bfun(double I, double kappa, double dhat) {
    double l = log(I);
    return kappa * dhat * dhat * (I - 1) * (I - 1) * l * l;
}
/* 2 b'(I) * ... the reference's flatten_pk1 scalar (RANK 2) */
static double pk1(double I, double kappa, double dhat) {
    double l = log(I);
    return 2 * (2 * kappa * dhat * dhat * l * (I - 1) * (I + I * l - 1)) / I;
}
/* d^2 b / ds^2, s = sqrt(I) (RANK 2 lambda0) */
static double lam0(double I, double kappa, double dhat) {
    double l = log(I);
    return -(4 * kappa * dhat * dhat *
             (4 * I + l - 3 * I * I * l * l + 6 * I * l - 2 * I * I + I * l * l - 7 * I * I * l - 2)) / I;
}

int or_barrier_pair(const double* x, const double* xr, const int q[4], double dhat, double kappa, int vid[4], int* cls,
                    double grad[12], double H[144], double* energy) {
    int a = q[0], b = q[1], c = q[2], w = q[3];
    int nv, moll = 0, dkind; /* dkind: 0 pp, 1 pe, 2 pt, 3 ee */
    int ea0 = 0, ea1 = 0, eb0 = 0, eb1 = 0; /* mollifier edges as vertex slots */
    memset(grad, 0, sizeof(double) * 12);
    memset(H, 0, sizeof(double) * 144);
    if (energy) *energy = 0.0;
    if (a >= 0) {
        if (w >= 0) { vid[0] = a; vid[1] = b; vid[2] = c; vid[3] = w; nv = 4; dkind = 3; *cls = 4; }
        else {
            vid[0] = a; vid[1] = b; vid[2] = c; vid[3] = -w - 1; nv = 4; dkind = 3; *cls = 4; moll = 1;
            ea0 = 0; ea1 = 1; eb0 = 2; eb1 = 3;
        }
    } else {
        int v0 = -a - 1;
        if (c < 0) {
            if (b < 0) {
                vid[0] = v0; vid[1] = -b - 1; vid[2] = -c - 1; vid[3] = -w - 1; nv = 4; dkind = 0; *cls = 4; moll = 1;
                ea0 = 0; ea1 = 2; eb0 = 1; eb1 = 3;
            } else { vid[0] = v0; vid[1] = b; nv = 2; dkind = 0; *cls = 2; }
        } else if (w < 0) {
            if (b < 0) {
                vid[0] = v0; vid[1] = -b - 1; vid[2] = c; vid[3] = -w - 1; nv = 4; dkind = 1; *cls = 4; moll = 1;
                ea0 = 0; ea1 = 3; eb0 = 1; eb1 = 2;
            } else { vid[0] = v0; vid[1] = b; vid[2] = c; nv = 3; dkind = 1; *cls = 3; }
        } else { vid[0] = v0; vid[1] = b; vid[2] = c; vid[3] = w; nv = 4; dkind = 2; *cls = 4; }
    }
    double gd[12] = {0};
    double d;
    switch (dkind) {
        case 0: d = d_pp(X(x, vid[0]), X(x, vid[1]), gd); break;
        case 1: d = d_pe(X(x, vid[0]), X(x, vid[1]), X(x, vid[2]), gd); break;
        case 2: d = d_pt(X(x, vid[0]), X(x, vid[1]), X(x, vid[2]), X(x, vid[3]), gd); break;
        default: d = d_ee(X(x, vid[0]), X(x, vid[1]), X(x, vid[2]), X(x, vid[3]), gd); break;
    }
    int nd = 3 * nv;
    double I = d / dhat, s = sqrt(I);
    double gs[12];
    for (int k = 0; k < 12; ++k) gs[k] = (k < nd && d > 0) ? gd[k] / (2 * sqrt(d * dhat)) : 0.0;
    if (!moll) {
        if (energy) *energy = bfun(I, kappa, dhat);
        double p = pk1(I, kappa, dhat) * s;
        double l0 = lam0((d < 1e-6 * dhat) ? 1e-6 : I, kappa, dhat);
        for (int k = 0; k < nd; ++k) grad[k] = gs[k] * p;
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j) H[nd * i + j] = l0 * gs[i] * gs[j];
        return nv;
    }
    /* mollified */
    int sl[4] = {ea0, ea1, eb0, eb1};
    double gc[12] = {0};
    double cc = cross_norm(X(x, vid[sl[0]]), X(x, vid[sl[1]]), X(x, vid[sl[2]]), X(x, vid[sl[3]]), sl[0], sl[1], sl[2],
                           sl[3], gc);
    double I1 = cc * cc;
    if (I1 == 0.0) return nv; /* reference returns without contribution */
    double r0[3], r1[3];
    sub3(X(xr, vid[sl[0]]), X(xr, vid[sl[1]]), r0);
    sub3(X(xr, vid[sl[2]]), X(xr, vid[sl[3]]), r1);
    double eps = 1e-3 * dot3(r0, r0) * dot3(r1, r1);
    double l = log(I), K2 = kappa * dhat * dhat;
    if (energy) *energy = -I1 * (I1 - 2 * eps) / (eps * eps) * bfun(I, kappa, dhat);
    double p1 = -kappa * 2 * (2 * dhat * dhat * l * l * (I1 - eps) * (I - 1) * (I - 1)) / (eps * eps);
    double p2 = -kappa * 2 * (2 * I1 * dhat * dhat * l * (I1 - 2 * eps) * (I - 1) * (I + I * l - 1)) / (I * eps * eps);
    for (int k = 0; k < 12; ++k) grad[k] = gc[k] * cc * p1 + gs[k] * s * p2;
    double l10 = -kappa * (4 * dhat * dhat * l * l * (I - 1) * (I - 1) * (3 * I1 - eps)) / (eps * eps);
    double l20 = kappa * (4 * I1 * dhat * dhat * (I1 - 2 * eps) *
                          (4 * I + l - 3 * I * I * l * l + 6 * I * l - 2 * I * I + I * l * l - 7 * I * I * l - 2)) /
                 (I * eps * eps);
    double lg = -kappa * 4 * cc * s * (4 * dhat * dhat * l * (I1 - eps) * (I - 1) * (I + I * l - 1)) / (I * eps * eps);
    (void)K2;
    /* PSD projection of the 2x2 (makePDGeneral<2>, S/GIPC.cu:31-57) */
    double m00 = l10, m01 = lg, m11 = l20;
    double tr = m00 + m11, det = m00 * m11 - m01 * m01;
    double disc = sqrt(fmax(0.25 * tr * tr - det, 0.0));
    double e0 = 0.5 * tr - disc, e1 = 0.5 * tr + disc;
    if (e0 < 0.0) {
        /* eigenvectors */
        double v0x, v0y;
        if (fabs(m01) > 0) { v0x = e0 - m11; v0y = m01; }
        else if (m00 <= m11) { v0x = 1; v0y = 0; }
        else { v0x = 0; v0y = 1; }
        double nv0 = sqrt(v0x * v0x + v0y * v0y);
        v0x /= nv0; v0y /= nv0;
        double v1x = -v0y, v1y = v0x;
        double pe = e1 > 0 ? e1 : 0.0;
        m00 = pe * v1x * v1x; m01 = pe * v1x * v1y; m11 = pe * v1y * v1y;
    }
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j)
            H[12 * i + j] = gc[i] * (m00 * gc[j] + m01 * gs[j]) + gs[i] * (m01 * gc[j] + m11 * gs[j]);
    return nv;
}

void or_barrier_assemble(int np, const int* pairs, const int* mi, const int counts[3], const double* x,
                         const double* xr, double dhat, double kappa, const int* btype, double* fb, double* vals,
                         int* rows, int* cols) {
    double g[12], H[144];
    int vid[4], cls;
    for (int p = 0; p < np; ++p) {
        int nv = or_barrier_pair(x, xr, pairs + 4 * p, dhat, kappa, vid, &cls, g, H, NULL);
        for (int a = 0; a < nv; ++a)
            for (int k = 0; k < 3; ++k) fb[3 * (size_t)vid[a] + k] += g[3 * a + k];
        size_t off;
        if (cls == 4) off = 16 * (size_t)mi[p];
        else if (cls == 3) off = 16 * (size_t)counts[2] + 9 * (size_t)mi[p];
        else off = 16 * (size_t)counts[2] + 9 * (size_t)counts[1] + 4 * (size_t)mi[p];
        int nd = 3 * nv, kk = 0;
        for (int ii = 0; ii < nv; ++ii)
            for (int jj = 0; jj < nv; ++jj, ++kk) {
                int r = vid[ii], c = vid[jj];
                int keep = r <= c;
                int zero = !keep || (btype && (btype[r] != 0 || btype[c] != 0));
                double* B = vals + 9 * (off + kk);
                for (int pp = 0; pp < 3; ++pp)
                    for (int qq = 0; qq < 3; ++qq) B[pp + 3 * qq] = zero ? 0.0 : H[nd * (3 * ii + pp) + 3 * jj + qq];
                rows[off + kk] = keep ? r : c;
                cols[off + kk] = keep ? c : r;
            }
    }
}

double or_barrier_energy(int np, const int* pairs, const double* x, const double* xr, double dhat, double kappa) {
    double g[12], H[144], e, tot = 0.0;
    int vid[4], cls;
    for (int p = 0; p < np; ++p) {
        or_barrier_pair(x, xr, pairs + 4 * p, dhat, kappa, vid, &cls, g, H, &e);
        tot += e;
    }
    return tot;
}

/* ---- classification (S/mlbvh.cu:288-532) and encoding (:535-1060) ------- */
static double det3c(const double* c0, const double* c1, const double* c2) {
    /* determinant of the matrix with columns c0, c1, c2 */
    double t[3];
    cross3(c1, c2, t);
    return dot3(c0, t);
}

static void pe_param(const double* v0, const double* a, const double* b, const double* nvec, double* px, double* py) {
    double b0[3], b1[3], b2[3];
    sub3(b, a, b0);
    cross3(b0, nvec, b1);
    sub3(v0, a, b2);
    double D = det3c(b0, b1, nvec);
    *px = det3c(b2, b1, nvec) / D;
    *py = det3c(b0, b2, nvec) / D;
}

static int dtype_pt(const double* v0, const double* v1, const double* v2, const double* v3) {
    double b0[3], b1[3], nvec[3];
    sub3(v2, v1, b0); sub3(v3, v1, b1); cross3(b0, b1, nvec);
    double p0x, p0y, p1x, p1y, p2x, p2y;
    pe_param(v0, v1, v2, nvec, &p0x, &p0y);
    if (p0x > 0 && p0x < 1 && p0y >= 0) return 3;
    pe_param(v0, v2, v3, nvec, &p1x, &p1y);
    if (p1x > 0.0 && p1x < 1.0 && p1y >= 0.0) return 4;
    pe_param(v0, v3, v1, nvec, &p2x, &p2y);
    if (p2x > 0.0 && p2x < 1.0 && p2y >= 0.0) return 5;
    if (p0x <= 0.0 && p2x >= 1.0) return 0;
    if (p1x <= 0.0 && p0x >= 1.0) return 1;
    if (p2x <= 0.0 && p1x >= 1.0) return 2;
    return 6;
}

static int dtype_ee(const double* v0, const double* v1, const double* v2, const double* v3) {
    double u[3], v[3], w[3], uv[3];
    sub3(v1, v0, u); sub3(v3, v2, v); sub3(v0, v2, w);
    double a = dot3(u, u), b = dot3(u, v), c = dot3(v, v), d = dot3(u, w), e = dot3(v, w);
    double D = a * c - b * b, tD = D, sN, tN;
    int def = 8;
    sN = (b * e - c * d);
    if (sN <= 0.0) { tN = e; tD = c; def = 2; }
    else if (sN >= D) { tN = e + b; tD = c; def = 5; }
    else {
        tN = (a * e - b * d);
        cross3(u, v, uv);
        if (tN > 0.0 && tN < tD && (dot3(w, uv) == 0.0 || dot3(uv, uv) < 1.0e-20 * a * c)) {
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

int or_classify(int npt, const int* pt, int nee, const int* ee, const double* x, const double* xr, double dhat,
                int* pairs, int* mi, int counts[3]) {
    int np = 0;
    counts[0] = counts[1] = counts[2] = 0;
#define EMIT(A, B, C, D, CLS)                        \
    do {                                             \
        pairs[4 * np] = (A); pairs[4 * np + 1] = (B);  \
        pairs[4 * np + 2] = (C); pairs[4 * np + 3] = (D); \
        mi[np] = counts[(CLS) - 2]++;                \
        np++;                                        \
    } while (0)
    for (int i = 0; i < npt; ++i) {
        int id0 = pt[4 * i], id1 = pt[4 * i + 1], id2 = pt[4 * i + 2], id3 = pt[4 * i + 3];
        const double *v0 = X(x, id0), *v1 = X(x, id1), *v2 = X(x, id2), *v3 = X(x, id3);
        int t = dtype_pt(v0, v1, v2, v3);
        double d;
        switch (t) {
            case 0: d = d_pp(v0, v1, NULL); if (d < dhat) EMIT(-id0 - 1, id1, -1, -1, 2); break;
            case 1: d = d_pp(v0, v2, NULL); if (d < dhat) EMIT(-id0 - 1, id2, -1, -1, 2); break;
            case 2: d = d_pp(v0, v3, NULL); if (d < dhat) EMIT(-id0 - 1, id3, -1, -1, 2); break;
            case 3: d = d_pe(v0, v1, v2, NULL); if (d < dhat) EMIT(-id0 - 1, id1, id2, -1, 3); break;
            case 4: d = d_pe(v0, v2, v3, NULL); if (d < dhat) EMIT(-id0 - 1, id2, id3, -1, 3); break;
            case 5: d = d_pe(v0, v3, v1, NULL); if (d < dhat) EMIT(-id0 - 1, id3, id1, -1, 3); break;
            case 6: d = d_pt(v0, v1, v2, v3, NULL); if (d < dhat) EMIT(-id0 - 1, id1, id2, id3, 4); break;
            default: break;
        }
    }
    for (int i = 0; i < nee; ++i) {
        int id0 = ee[4 * i], id1 = ee[4 * i + 1], id2 = ee[4 * i + 2], id3 = ee[4 * i + 3];
        const double *v0 = X(x, id0), *v1 = X(x, id1), *v2 = X(x, id2), *v3 = X(x, id3);
        int t = dtype_ee(v0, v1, v2, v3);
        double r0[3], r1[3], u[3], v[3], cr[3];
        sub3(X(xr, id0), X(xr, id1), r0); sub3(X(xr, id2), X(xr, id3), r1);
        double eps = 1e-3 * dot3(r0, r0) * dot3(r1, r1);
        sub3(v0, v1, u); sub3(v2, v3, v); cross3(u, v, cr);
        int par = dot3(cr, cr) < eps;
        int add_e = par ? -i - 2 : -1;  /* obj_idx = candidate index */
        double d;
        switch (t) {
            case 0: d = d_pp(v0, v2, NULL); if (d < dhat) EMIT(-id0 - 1, id2, -1, add_e, 2); break;
            case 1: d = d_pp(v0, v3, NULL); if (d < dhat) EMIT(-id0 - 1, id3, -1, add_e, 2); break;
            case 2: d = d_pe(v0, v2, v3, NULL); if (d < dhat) EMIT(-id0 - 1, id2, id3, add_e, 3); break;
            case 3: d = d_pp(v1, v2, NULL); if (d < dhat) EMIT(-id1 - 1, id2, -1, add_e, 2); break;
            case 4: d = d_pp(v1, v3, NULL); if (d < dhat) EMIT(-id1 - 1, id3, -1, add_e, 2); break;
            case 5: d = d_pe(v1, v2, v3, NULL); if (d < dhat) EMIT(-id1 - 1, id2, id3, add_e, 3); break;
            case 6: d = d_pe(v2, v0, v1, NULL); if (d < dhat) EMIT(-id2 - 1, id0, id1, add_e, 3); break;
            case 7: d = d_pe(v3, v0, v1, NULL); if (d < dhat) EMIT(-id3 - 1, id0, id1, add_e, 3); break;
            case 8: d = d_ee(v0, v1, v2, v3, NULL); if (d < dhat) EMIT(id0, id1, id2, id3, 4); break;
            default: break;
        }
    }
#undef EMIT
    return np;
}

/* ======================================================================== */
/* Additive CCD (Li et al. 2021, "Codimensional incremental potential       */
/* contact"): ACCD.cu point_triangle_ccd :456-534, edge_edge_ccd :373-454,   */
/* thickness 0; step bound as GIPC::self_largestFeasibleStepSize            */
/* (GIPC.cu:8880-8915) with eta = 1 - slackness.                            */
/* ======================================================================== */
static double ccd_d2_pt(const double* p, const double* a, const double* b, const double* c) {
    switch (dtype_pt(p, a, b, c)) {
        case 0: return d_pp(p, a, NULL);
        case 1: return d_pp(p, b, NULL);
        case 2: return d_pp(p, c, NULL);
        case 3: return d_pe(p, a, b, NULL);
        case 4: return d_pe(p, b, c, NULL);
        case 5: return d_pe(p, c, a, NULL);
        default: return d_pt(p, a, b, c, NULL);
    }
}
static double ccd_d2_ee(const double* a0, const double* a1, const double* b0, const double* b1) {
    double d;
    switch (dtype_ee(a0, a1, b0, b1)) {
        case 0: d = d_pp(a0, b0, NULL); break;
        case 1: d = d_pp(a0, b1, NULL); break;
        case 2: d = d_pe(a0, b0, b1, NULL); break;
        case 3: d = d_pp(a1, b0, NULL); break;
        case 4: d = d_pp(a1, b1, NULL); break;
        case 5: d = d_pe(a1, b0, b1, NULL); break;
        case 6: d = d_pe(b0, a0, a1, NULL); break;
        case 7: d = d_pe(b1, a0, a1, NULL); break;
        default: d = d_ee(a0, a1, b0, b1, NULL); break;
    }
    if (d <= 0.0) {
        double m = d_pp(a0, b0, NULL), t;
        t = d_pp(a0, b1, NULL); if (t < m) m = t;
        t = d_pp(a1, b0, NULL); if (t < m) m = t;
        t = d_pp(a1, b1, NULL); if (t < m) m = t;
        d = m;
    }
    return d;
}

/* v: 4 vertices (12), u: their displacements (12); pt != 0: point-triangle */
double or_accd(int pt, const double* v_in, const double* u_in, double eta) {
    double v[12], u[12], mov[3];
    for (int k = 0; k < 3; ++k) mov[k] = -0.25 * (u_in[k] + u_in[3 + k] + u_in[6 + k] + u_in[9 + k]);
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 3; ++k) { v[3 * i + k] = v_in[3 * i + k]; u[3 * i + k] = u_in[3 * i + k] + mov[k]; }
    double n0 = dot3(u, u), n1 = dot3(u + 3, u + 3), n2 = dot3(u + 6, u + 6), n3 = dot3(u + 9, u + 9);
    double maxd;
    if (pt) maxd = sqrt(n0) + sqrt(fmax(n1, fmax(n2, n3)));
    else maxd = sqrt(fmax(n0, n1)) + sqrt(fmax(n2, n3));
    if (maxd == 0.0) return 1.0;
    double dd = pt ? ccd_d2_pt(v, v + 3, v + 6, v + 9) : ccd_d2_ee(v, v + 3, v + 6, v + 9);
    double dist = sqrt(dd), gap = eta * dd / dist, toc = 0.0;
    for (int it = 0; it < 50000; ++it) {
        double lb = (1.0 - eta) * dd / (dist * maxd);
        for (int i = 0; i < 12; ++i) v[i] += lb * u[i];
        dd = pt ? ccd_d2_pt(v, v + 3, v + 6, v + 9) : ccd_d2_ee(v, v + 3, v + 6, v + 9);
        dist = sqrt(dd);
        if (toc != 0.0 && dd / dist < gap) return toc;
        toc += lb;
        if (toc > 1.0) return 1.0;
    }
    return toc;
}

/* min over candidates of the time of impact along -dir, at most 1 */
double or_ccd_step_bound(int npt, const int* pt, int nee, const int* ee, const double* x, const double* dir,
                         double slack) {
    double t = 1.0, v[12], u[12];
    for (int i = 0; i < npt + nee; ++i) {
        const int* q = i < npt ? pt + 4 * i : ee + 4 * (i - npt);
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 3; ++k) { v[3 * a + k] = x[3 * (size_t)q[a] + k]; u[3 * a + k] = -dir[3 * (size_t)q[a] + k]; }
        double s = or_accd(i < npt, v, u, 1.0 - slack);
        if (s < t) t = s;
    }
    return t;
}
