/*
 * sgipc_oracle.c -- CPU restatement (FP64, plain C99) of Stiff-GIPC's
 * Newton-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP product path is
 * compared against (tests/, __graft_entry__.smoke) and the CPU baseline that
 * bench.py times beside the GPU.  Nothing in the product links or calls it.
 *
 * Parity pins (what each function is checked against, see tests/):
 *   or_svd3          reconstruction U S V^T = A, det U = det V = +1, ordering
 *   or_syev3         A Q = Q diag(w), Q orthonormal
 *   or_snk1_*        closed-form energy; gradient vs central FD of the energy;
 *                    Hessian vs numpy eigh-projection of the exact Hessian
 *   or_convert       brute-force dictionary accumulate (bit-exact indices)
 *   or_spmv_sym      dense numpy matvec of the expanded symmetric matrix
 *   or_mas_*         tables: brute-force connected components per bank;
 *                    apply: explicit dense R_l / D_l^{-1} construction
 *   or_pcg           scipy direct solve of the same system
 *   or_ground_*      energy closed form; gradient vs central FD of the
 *                    energy; Hessian vs central FD of the gradient
 *   or_fem_param_adjoint  central FD of the assembled gradient in mu, lam
 *   or_accd          head-on approach: (1 - eta) of the exact time of impact;
 *                    no penetration over [0, toc] on random motions
 *   exact SNK1 mode  (projection off) Hessian vs central FD of the gradient
 */
#include "sgipc_oracle_internal.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================== */
/* Givens rotations and the implicit-shift QR SVD of a 3x3 matrix.          */
/* Published method: Gast, Jiang, Teran, "Implicit-shifted symmetric QR     */
/* singular value decomposition of 3x3 matrices", UCLA tech. note (2016),   */
/* as used by S/QRSVD.hpp:219-317 and S/givens.hpp:16-213.                  */
/* ======================================================================== */
typedef struct { int i, k; double c, s; } giv_t;

static giv_t giv(int i, int k) { giv_t g = {i, k, 1.0, 0.0}; return g; }
/* [c -s; s c]^T-style rotation chosen so that (a,b) -> (*, 0) */
static void giv_conv(giv_t* g, double a, double b) {
    double d = a * a + b * b;
    g->c = 1.0; g->s = 0.0;
    if (d != 0.0) { double t = 1.0 / sqrt(d); g->c = a * t; g->s = -b * t; }
}
/* (a,b) -> (0, *) */
static void giv_unconv(giv_t* g, double a, double b) {
    double d = a * a + b * b;
    g->c = 0.0; g->s = 1.0;
    if (d != 0.0) { double t = 1.0 / sqrt(d); g->s = a * t; g->c = b * t; }
}
/* A <- G^T A on rows i,k (3x3 row-major) */
static void giv_rows(const giv_t* g, double* A) {
    for (int j = 0; j < 3; ++j) {
        double t1 = A[3 * g->i + j], t2 = A[3 * g->k + j];
        A[3 * g->i + j] = g->c * t1 - g->s * t2;
        A[3 * g->k + j] = g->s * t1 + g->c * t2;
    }
}
/* A <- A G on cols i,k */
static void giv_cols(const giv_t* g, double* A) {
    for (int r = 0; r < 3; ++r) {
        double t1 = A[3 * r + g->i], t2 = A[3 * r + g->k];
        A[3 * r + g->i] = g->c * t1 - g->s * t2;
        A[3 * r + g->k] = g->s * t1 + g->c * t2;
    }
}

static void zero_chase(double* H, double* U, double* V) {
    giv_t r1 = giv(0, 1), r2 = giv(1, 2), r3 = giv(1, 2);
    giv_conv(&r1, H[0], H[3]);
    if (H[3] != 0.0)
        giv_conv(&r2, H[0] * H[1] + H[3] * H[4], H[0] * H[2] + H[3] * H[5]);
    else
        giv_conv(&r2, H[1], H[2]);
    giv_rows(&r1, H);
    giv_cols(&r2, H);
    giv_cols(&r2, V);
    giv_conv(&r3, H[4], H[7]);
    giv_rows(&r3, H);
    giv_cols(&r1, U);
    giv_cols(&r3, U);
}

static void bidiag(double* H, double* U, double* V) {
    giv_t r = giv(1, 2);
    giv_conv(&r, H[3], H[6]);
    giv_rows(&r, H);
    giv_cols(&r, U);
    zero_chase(H, U, V);
}

/* 2x2 SVD of the block at (t,t): polar decomposition then symmetric Jacobi. */
static void svd2(double a00, double a01, double a10, double a11, double S[2],
                 giv_t* u, giv_t* v) {
    double x0 = a00 + a11, x1 = a10 - a01, d = sqrt(x0 * x0 + x1 * x1);
    if (d != 0.0) { u->c = x0 / d; u->s = -x1 / d; } else { u->c = 1.0; u->s = 0.0; }
    /* S_sym = G^T A */
    double s00 = u->c * a00 - u->s * a10, s01 = u->c * a01 - u->s * a11;
    double s11 = u->s * a01 + u->c * a11;
    double x = s00, y = s01, z = s11, y2 = y * y, cs, sn;
    if (y2 == 0.0) { cs = 1.0; sn = 0.0; S[0] = x; S[1] = z; }
    else {
        double tau = 0.5 * (x - z), w = sqrt(tau * tau + y2), t;
        t = (tau > 0.0) ? y / (tau + w) : y / (tau - w);
        cs = 1.0 / sqrt(t * t + 1.0);
        sn = -t * cs;
        double c2 = cs * cs, cs2y = 2.0 * cs * sn * y, s2 = sn * sn;
        S[0] = c2 * x - cs2y + s2 * z;
        S[1] = s2 * x + cs2y + c2 * z;
    }
    if (S[0] < S[1]) { double tt = S[0]; S[0] = S[1]; S[1] = tt; v->c = -sn; v->s = cs; }
    else { v->c = cs; v->s = sn; }
    /* u *= v */
    double nc = u->c * v->c - u->s * v->s, ns = u->s * v->c + u->c * v->s;
    u->c = nc; u->s = ns;
}

static void process_t(int t, const double* B, double* U, double S[3], double* V) {
    int other = (t == 1) ? 0 : 2;
    S[other] = B[4 * other];
    giv_t u = giv(t, t + 1), v = giv(t, t + 1);
    double s2[2];
    svd2(B[3 * t + t], B[3 * t + t + 1], B[3 * (t + 1) + t], B[3 * (t + 1) + t + 1], s2, &u, &v);
    S[t] = s2[0]; S[t + 1] = s2[1];
    giv_cols(&u, U);
    giv_cols(&v, V);
}

static void flip_col(double* M, int j) { for (int r = 0; r < 3; ++r) M[3 * r + j] = -M[3 * r + j]; }
static void swap_col(double* M, int a, int b) {
    for (int r = 0; r < 3; ++r) { double t = M[3 * r + a]; M[3 * r + a] = M[3 * r + b]; M[3 * r + b] = t; }
}
static void flip_sign(int j, double* U, double* S) { S[j] = -S[j]; flip_col(U, j); }
static void swapd(double* a, double* b) { double t = *a; *a = *b; *b = t; }

static void sort_sigma0(double* U, double* S, double* V) {
    if (fabs(S[1]) >= fabs(S[2])) {
        if (S[1] < 0) { flip_sign(1, U, S); flip_sign(2, U, S); }
        return;
    }
    if (S[2] < 0) { flip_sign(1, U, S); flip_sign(2, U, S); }
    swapd(&S[1], &S[2]); swap_col(U, 1, 2); swap_col(V, 1, 2);
    if (S[1] > S[0]) { swapd(&S[0], &S[1]); swap_col(U, 0, 1); swap_col(V, 0, 1); }
    else { flip_col(U, 2); flip_col(V, 2); }
}
static void sort_sigma1(double* U, double* S, double* V) {
    if (fabs(S[0]) >= S[1]) {
        if (S[0] < 0) { flip_sign(0, U, S); flip_sign(2, U, S); }
        return;
    }
    swapd(&S[0], &S[1]); swap_col(U, 0, 1); swap_col(V, 0, 1);
    if (fabs(S[1]) < fabs(S[2])) { swapd(&S[1], &S[2]); swap_col(U, 1, 2); swap_col(V, 1, 2); }
    else { flip_col(U, 1); flip_col(V, 1); }
    if (S[1] < 0) { flip_sign(1, U, S); flip_sign(2, U, S); }
}

static double wilkinson(double a1, double b1, double a2) {
    double d = 0.5 * (a1 - a2), bs = b1 * b1;
    return a2 - copysign(bs / (fabs(d) + sqrt(d * d + bs)), d);
}

void or_svd3(const double A[9], double U[9], double S[3], double V[9]) {
    double B[9];
    memcpy(B, A, sizeof(B));
    for (int i = 0; i < 9; ++i) { U[i] = (i % 4 == 0); V[i] = (i % 4 == 0); }
    bidiag(B, U, V);
    giv_t r = giv(0, 1);
    double al[3] = {B[0], B[4], B[8]}, be[2] = {B[1], B[5]};
    double ga[2] = {al[0] * be[0], al[1] * be[1]};
    double eta = DBL_EPSILON * 128.0;
    double nrm = 0.5 * sqrt(al[0] * al[0] + al[1] * al[1] + al[2] * al[2] + be[0] * be[0] + be[1] * be[1]);
    double tol = eta * (nrm > 1.0 ? nrm : 1.0);
    while (fabs(al[0]) > tol && fabs(al[1]) > tol && fabs(al[2]) > tol && fabs(be[0]) > tol &&
           fabs(be[1]) > tol) {
        double mu = wilkinson(al[1] * al[1] + be[0] * be[0], ga[1], al[2] * al[2] + be[1] * be[1]);
        giv_conv(&r, al[0] * al[0] - mu, ga[0]);
        giv_cols(&r, B);
        giv_cols(&r, V);
        zero_chase(B, U, V);
        al[0] = B[0]; al[1] = B[4]; al[2] = B[8]; be[0] = B[1]; be[1] = B[5];
        ga[0] = al[0] * be[0]; ga[1] = al[1] * be[1];
    }
    if (fabs(be[1]) <= tol) {
        process_t(0, B, U, S, V); sort_sigma0(U, S, V);
    } else if (fabs(be[0]) <= tol) {
        process_t(1, B, U, S, V); sort_sigma1(U, S, V);
    } else if (fabs(al[1]) <= tol) {
        giv_t r1 = giv(1, 2);
        giv_unconv(&r1, B[5], B[8]);
        giv_rows(&r1, B); giv_cols(&r1, U);
        process_t(0, B, U, S, V); sort_sigma0(U, S, V);
    } else if (fabs(al[2]) <= tol) {
        giv_t r1 = giv(1, 2), r2 = giv(0, 2);
        giv_conv(&r1, B[4], B[5]);
        giv_cols(&r1, B); giv_cols(&r1, V);
        giv_conv(&r2, B[0], B[2]);
        giv_cols(&r2, B); giv_cols(&r2, V);
        process_t(0, B, U, S, V); sort_sigma0(U, S, V);
    } else if (fabs(al[0]) <= tol) {
        giv_t r1 = giv(0, 1), r2 = giv(0, 2);
        giv_unconv(&r1, B[1], B[4]);
        giv_rows(&r1, B); giv_cols(&r1, U);
        giv_unconv(&r2, B[2], B[8]);
        giv_rows(&r2, B); giv_cols(&r2, U);
        process_t(1, B, U, S, V); sort_sigma1(U, S, V);
    }
}

/* Cyclic Jacobi for a symmetric 3x3 matrix; ascending eigenvalues. */
void or_syev3(const double A[9], double w[3], double Q[9]) {
    double a[9];
    memcpy(a, A, sizeof(a));
    for (int i = 0; i < 9; ++i) Q[i] = (i % 4 == 0);
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
        double dg = a[0] * a[0] + a[4] * a[4] + a[8] * a[8];
        if (off <= 1e-36 * dg || off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double apq = a[3 * p + q];
                if (apq == 0.0) continue;
                double theta = (a[3 * q + q] - a[3 * p + p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) { /* A <- A J */
                    double akp = a[3 * k + p], akq = a[3 * k + q];
                    a[3 * k + p] = c * akp - s * akq;
                    a[3 * k + q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) { /* A <- J^T A */
                    double apk = a[3 * p + k], aqk = a[3 * q + k];
                    a[3 * p + k] = c * apk - s * aqk;
                    a[3 * q + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    double qkp = Q[3 * k + p], qkq = Q[3 * k + q];
                    Q[3 * k + p] = c * qkp - s * qkq;
                    Q[3 * k + q] = s * qkp + c * qkq;
                }
            }
    }
    w[0] = a[0]; w[1] = a[4]; w[2] = a[8];
    for (int i = 0; i < 2; ++i) /* ascending selection sort with columns */
        for (int j = i + 1; j < 3; ++j)
            if (w[j] < w[i]) { swapd(&w[i], &w[j]); swap_col(Q, i, j); }
}

/* ======================================================================== */
/* SNK1 tet energy, gradient and projected Hessian                          */
/* ======================================================================== */
static double det3(const double* F) {
    return F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
           F[2] * (F[3] * F[7] - F[4] * F[6]);
}

/* F = Ds DmInv, Ds columns x1-x0, x2-x0, x3-x0 (S/femEnergy.cu:324-350) */
static void tet_F(const double xt[12], const double DmInv[9], double F[9]) {
    double Ds[9];
    for (int i = 0; i < 3; ++i)
        for (int a = 0; a < 3; ++a) Ds[3 * i + a] = xt[3 * (a + 1) + i] - xt[i];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += Ds[3 * i + k] * DmInv[3 * k + j];
            F[3 * i + j] = s;
        }
}

/* Psi = 0.5 vol [mu (I2-3) + lam (J - 1 - mu/lam)^2]  (S/femEnergy.cu:1032-1050) */
double or_snk1_energy(const double xt[12], const double DmInv[9], double vol, double mu,
                      double lam) {
    double F[9];
    tet_F(xt, DmInv, F);
    double I2 = 0;
    for (int i = 0; i < 9; ++i) I2 += F[i] * F[i];
    double Jm = det3(F) - 1.0 - mu / lam;
    return 0.5 * (mu * (I2 - 3.0) + lam * Jm * Jm) * vol;
}

/* Bm (4x3): dF(i,j)/dx_{a,k} = delta_ik Bm(a,j) (S/femEnergy.cu:1389-1436) */
static void tet_Bm(const double DmInv[9], double Bm[12]) {
    for (int j = 0; j < 3; ++j) {
        Bm[j] = -(DmInv[j] + DmInv[3 + j] + DmInv[6 + j]);
        for (int a = 0; a < 3; ++a) Bm[3 * (a + 1) + j] = DmInv[3 * a + j];
    }
}

/* 1: clamp negative modes (the reference's PSD Newton Hessian); 0: exact
 * second derivative (all nine modes), as the library's adjoint mode */
int or_fem_project = 1;
void or_set_fem_projection(int on) { or_fem_project = on ? 1 : 0; }

/* Analytic eigensystem of the SNK1 dPsi/dF^2 with negative modes clamped
 * (S/femEnergy.cu:1593-1717).  vec index = i + 3 j (column-major). */
void or_snk1_hq(const double F[9], double mu, double lam, double H81[81]) {
    double U[9], S[3], V[9];
    or_svd3(F, U, S, V);
    double J = det3(F);
    double ev[9], Q[9][9]; /* Q[k] = eigenvector k as vec9 */
    double sc = lam * (J - 1.0) - mu;
    for (int i = 0; i < 3; ++i) { ev[i] = mu + S[i] * sc; ev[3 + i] = mu - S[i] * sc; }
    const double h = 1.0 / sqrt(2.0);
    /* twist/flip: pairs (p,q) of singular directions for eigen index i */
    const int pq[3][2] = {{1, 2}, {0, 2}, {0, 1}};
    for (int t = 0; t < 3; ++t) {
        int p = pq[t][0], q = pq[t][1];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double a = h * U[3 * r + p] * V[3 * c + q]; /* u_p v_q^T */
                double b = h * U[3 * r + q] * V[3 * c + p]; /* u_q v_p^T */
                Q[t][r + 3 * c] = b - a;
                Q[3 + t][r + 3 * c] = a + b;
            }
    }
    double s00 = S[0] * S[0], s11 = S[1] * S[1], s22 = S[2] * S[2];
    double Am[9], w[3], E[9];
    double sc2 = lam * (2.0 * J - 1.0) - mu;
    Am[0] = mu + lam * s11 * s22; Am[4] = mu + lam * s00 * s22; Am[8] = mu + lam * s00 * s11;
    Am[1] = Am[3] = sc2 * S[2];
    Am[2] = Am[6] = sc2 * S[1];
    Am[5] = Am[7] = sc2 * S[0];
    or_syev3(Am, w, E);
    for (int k = 0; k < 3; ++k) {
        ev[6 + k] = w[k];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = 0;
                for (int m = 0; m < 3; ++m) s += U[3 * r + m] * E[3 * m + k] * V[3 * c + m];
                Q[6 + k][r + 3 * c] = s;
            }
    }
    if (or_fem_project) for (int k = 0; k < 9; ++k) if (ev[k] < 0.0) ev[k] = 0.0;
    for (int a = 0; a < 9; ++a)
        for (int b = 0; b < 9; ++b) {
            double s = 0;
            for (int k = 0; k < 9; ++k) s += ev[k] * Q[k][a] * Q[k][b];
            H81[9 * a + b] = s;
        }
}

void or_snk1_tet(const double xt[12], const double DmInv[9], double vol, double mu, double lam,
                 double dt, double grad12[12], double H144[144]) {
    double F[9], Bm[12], P[9], cof[9];
    tet_F(xt, DmInv, F);
    tet_Bm(DmInv, Bm);
    double J = det3(F);
    /* cofactor dJ/dF (S/femEnergy.cu:432-459) */
    cof[0] = F[4] * F[8] - F[5] * F[7]; cof[1] = F[5] * F[6] - F[3] * F[8]; cof[2] = F[3] * F[7] - F[4] * F[6];
    cof[3] = F[7] * F[2] - F[8] * F[1]; cof[4] = F[8] * F[0] - F[6] * F[2]; cof[5] = F[6] * F[1] - F[7] * F[0];
    cof[6] = F[1] * F[5] - F[4] * F[2]; cof[7] = F[2] * F[3] - F[0] * F[5]; cof[8] = F[0] * F[4] - F[1] * F[3];
    double k = lam * (J - 1.0 - mu / lam);
    for (int i = 0; i < 9; ++i) P[i] = mu * F[i] + k * cof[i];
    double s = dt * dt * vol;
    for (int a = 0; a < 4; ++a)
        for (int kk = 0; kk < 3; ++kk) {
            double g = 0;
            for (int j = 0; j < 3; ++j) g += Bm[3 * a + j] * P[3 * kk + j];
            grad12[3 * a + kk] = s * g;
        }
    double Hq[81];
    or_snk1_hq(F, mu, lam, Hq);
    for (int a = 0; a < 4; ++a)
        for (int kk = 0; kk < 3; ++kk)
            for (int b = 0; b < 4; ++b)
                for (int l = 0; l < 3; ++l) {
                    double h = 0;
                    for (int j = 0; j < 3; ++j)
                        for (int jp = 0; jp < 3; ++jp)
                            h += Bm[3 * a + j] * Bm[3 * b + jp] * Hq[9 * (kk + 3 * j) + (l + 3 * jp)];
                    H144[12 * (3 * a + kk) + 3 * b + l] = s * h;
                }
}

void or_fem_precompute(int n, int T, const double* xr, const uint32_t* tets, double E, double nu,
                       double density, double* DmInv, double* vol, double* mass, double* mu,
                       double* lam) {
    memset(mass, 0, sizeof(double) * (size_t)n);
    double lmu = E / (2.0 * (1.0 + nu));
    double llam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    for (int t = 0; t < T; ++t) {
        const uint32_t* v = tets + 4 * t;
        double Dm[9];
        for (int i = 0; i < 3; ++i)
            for (int a = 0; a < 3; ++a) Dm[3 * i + a] = xr[3 * v[a + 1] + i] - xr[3 * v[0] + i];
        double d = det3(Dm), id = 1.0 / d, *M = DmInv + 9 * t;
        M[0] = (Dm[4] * Dm[8] - Dm[5] * Dm[7]) * id; M[1] = (Dm[2] * Dm[7] - Dm[1] * Dm[8]) * id;
        M[2] = (Dm[1] * Dm[5] - Dm[2] * Dm[4]) * id; M[3] = (Dm[5] * Dm[6] - Dm[3] * Dm[8]) * id;
        M[4] = (Dm[0] * Dm[8] - Dm[2] * Dm[6]) * id; M[5] = (Dm[2] * Dm[3] - Dm[0] * Dm[5]) * id;
        M[6] = (Dm[3] * Dm[7] - Dm[4] * Dm[6]) * id; M[7] = (Dm[1] * Dm[6] - Dm[0] * Dm[7]) * id;
        M[8] = (Dm[0] * Dm[4] - Dm[1] * Dm[3]) * id;
        vol[t] = fabs(d) / 6.0;
        for (int a = 0; a < 4; ++a) mass[v[a]] += vol[t] * density / 4.0;
        mu[t] = 4.0 * lmu / 3.0;
        lam[t] = llam + 5.0 * lmu / 6.0;
    }
}

void or_fem_assemble(int n, int T, const double* x, const uint32_t* tets, const double* DmInv,
                     const double* vol, const double* mu, const double* lam, double dt,
                     const int* btype, int off, double* grad, double* vals, int* rows, int* cols) {
    if (or_threads > 1) {
        or_fem_assemble_mt(n, T, x, tets, DmInv, vol, mu, lam, dt, btype, off, grad, vals, rows, cols);
        return;
    }
    double xt[12], g[12], H[144];
    for (int t = 0; t < T; ++t) {
        const uint32_t* v = tets + 4 * t;
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 3; ++k) xt[3 * a + k] = x[3 * v[a] + k];
        or_snk1_tet(xt, DmInv + 9 * t, vol[t], mu[t], lam[t], dt, g, H);
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 3; ++k) grad[3 * v[a] + k] += g[3 * a + k];
        int kk = 0;
        for (int ii = 0; ii < 4; ++ii)
            for (int jj = ii; jj < 4; ++jj, ++kk) {
                int r = (int)v[ii] + off, c = (int)v[jj] + off, tr = 0;
                if (r > c) { int s = r; r = c; c = s; tr = 1; }
                int zero = btype && (btype[r - off] != 0 || btype[c - off] != 0);
                double* B = vals + 9 * (10 * (size_t)t + kk);
                for (int p = 0; p < 3; ++p)
                    for (int q = 0; q < 3; ++q) {
                        double h = tr ? H[12 * (3 * jj + p) + 3 * ii + q] : H[12 * (3 * ii + p) + 3 * jj + q];
                        B[p + 3 * q] = zero ? 0.0 : h;
                    }
                rows[10 * (size_t)t + kk] = r;
                cols[10 * (size_t)t + kk] = c;
            }
    }
}

void or_kinetic_gradient(int n, const double* x, const double* xt, const double* mass, double* g) {
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) g[3 * i + k] = (x[3 * i + k] - xt[3 * i + k]) * mass[i];
}

void or_mass_triplets(int n, const double* mass, int off, double* vals, int* rows, int* cols) {
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 9; ++k) vals[9 * (size_t)i + k] = (k % 4 == 0) ? mass[i] : 0.0;
        rows[i] = cols[i] = i + off;
    }
}

void or_fem_rhs(int n, const double* fb, const double* sh, const int* btype, double* b) {
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k)
            b[3 * i + k] = (btype && btype[i] != 0) ? 0.0 : fb[3 * i + k] + sh[3 * i + k];
}

/* ======================================================================== */
/* Triplet -> BCOO                                                          */
/* ======================================================================== */
typedef struct { uint64_t key; int t; } kt_t;
static int kt_cmp(const void* a, const void* b) {
    const kt_t *x = (const kt_t*)a, *y = (const kt_t*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return (x->t > y->t) - (x->t < y->t);
}

int or_convert(int nt, const int* rows, const int* cols, const double* vals, int* orow, int* ocol,
               double* oval) {
    if (nt < 1) return 0;
    if (or_threads > 1) return or_convert_mt(nt, rows, cols, vals, orow, ocol, oval);
    kt_t* e = (kt_t*)malloc(sizeof(kt_t) * (size_t)nt);
    for (int i = 0; i < nt; ++i) {
        e[i].key = ((uint64_t)(uint32_t)rows[i] << 32) + (uint64_t)(uint32_t)cols[i];
        e[i].t = i;
    }
    qsort(e, (size_t)nt, sizeof(kt_t), kt_cmp);
    int K = -1;
    uint64_t prev = 0;
    for (int i = 0; i < nt; ++i) {
        const double* v = vals + 9 * (size_t)e[i].t;
        if (i == 0 || e[i].key != prev) {
            ++K;
            orow[K] = (int)(e[i].key >> 32);
            ocol[K] = (int)(e[i].key & 0xffffffffu);
            for (int k = 0; k < 9; ++k) oval[9 * (size_t)K + k] = v[k];
            prev = e[i].key;
        } else {
            for (int k = 0; k < 9; ++k) oval[9 * (size_t)K + k] += v[k];
        }
    }
    free(e);
    return K + 1;
}

void or_spmv_sym(int K, const int* rows, const int* cols, const double* vals, int nb, double a,
                 const double* x, double b, double* y) {
    for (int i = 0; i < 3 * nb; ++i) y[i] = (b != 0.0) ? b * y[i] : 0.0;
    for (int k = 0; k < K; ++k) {
        int i = rows[k], j = cols[k];
        const double* M = vals + 9 * (size_t)k; /* column-major */
        for (int p = 0; p < 3; ++p) {
            double s = 0;
            for (int q = 0; q < 3; ++q) s += M[p + 3 * q] * x[3 * j + q];
            y[3 * i + p] += a * s;
        }
        if (i != j)
            for (int q = 0; q < 3; ++q) {
                double s = 0;
                for (int p = 0; p < 3; ++p) s += M[p + 3 * q] * x[3 * i + p];
                y[3 * j + q] += a * s;
            }
    }
}

/* ======================================================================== */
/* MAS inputs                                                               */
/* ======================================================================== */
typedef struct { uint32_t a, b; } edge_t;
typedef struct { uint32_t lo, hi; size_t pos; } etag_t;
static int etag_cmp(const void* p, const void* q) {
    const etag_t *x = (const etag_t*)p, *y = (const etag_t*)q;
    if (x->lo != y->lo) return x->lo < y->lo ? -1 : 1;
    if (x->hi != y->hi) return x->hi < y->hi ? -1 : 1;
    return (x->pos > y->pos) - (x->pos < y->pos);
}
static int edge_cmp_ab(const void* p, const void* q) {
    const edge_t *x = (const edge_t*)p, *y = (const edge_t*)q;
    if (x->a != y->a) return x->a < y->a ? -1 : 1;
    return (x->b > y->b) - (x->b < y->b);
}

/* S/load_mesh.cpp:934-1040: tet edges deduplicated regardless of orientation,
 * keeping the first-seen orientation (a,b) (tet order, then pair order
 * (0,1),(0,2),(0,3),(1,2),(1,3),(2,3)); edges iterated in (a,b) order; each
 * edge appends b to a's list and a to b's list. */
int or_vertex_neighbors(int n, int T, const uint32_t* tets, uint32_t* start, uint32_t* num,
                        uint32_t* list) {
    size_t ne = 6 * (size_t)T;
    edge_t* e = (edge_t*)malloc(sizeof(edge_t) * (ne + 1));
    etag_t* tg = (etag_t*)malloc(sizeof(etag_t) * (ne + 1));
    const int pr[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    size_t q = 0;
    for (int t = 0; t < T; ++t)
        for (int p = 0; p < 6; ++p, ++q) {
            uint32_t a = tets[4 * t + pr[p][0]], b = tets[4 * t + pr[p][1]];
            e[q].a = a; e[q].b = b;
            tg[q].lo = a < b ? a : b; tg[q].hi = a < b ? b : a; tg[q].pos = q;
        }
    qsort(tg, ne, sizeof(etag_t), etag_cmp);
    edge_t* u = (edge_t*)malloc(sizeof(edge_t) * (ne + 1));
    size_t nu = 0;
    for (size_t i = 0; i < ne; ++i) {
        if (i > 0 && tg[i].lo == tg[i - 1].lo && tg[i].hi == tg[i - 1].hi) continue;
        u[nu++] = e[tg[i].pos];
    }
    qsort(u, nu, sizeof(edge_t), edge_cmp_ab);
    memset(num, 0, sizeof(uint32_t) * (size_t)n);
    for (size_t i = 0; i < nu; ++i) { num[u[i].a]++; num[u[i].b]++; }
    uint32_t off = 0;
    for (int i = 0; i < n; ++i) { start[i] = off; off += num[i]; }
    uint32_t* cur = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
    for (size_t i = 0; i < nu; ++i) {
        list[start[u[i].a] + cur[u[i].a]++] = u[i].b;
        list[start[u[i].b] + cur[u[i].b]++] = u[i].a;
    }
    free(cur); free(u); free(tg); free(e);
    return (int)off;
}

/* S/gl_main.cu:1066-1093 (setMAS_partition); part ids ascending. */
int or_bank_map(int n, const int* part_id, int bank, int* part_map, int* real_map) {
    if (n < 1) return 0;
    int nparts = part_id[n - 1] + 1, M = bank * nparts, rank = 0;
    for (int i = 0; i < M; ++i) part_map[i] = -1;
    for (int i = 0; i < n; ++i) {
        if (rank >= bank) return -1; /* part larger than a bank */
        part_map[bank * part_id[i] + rank] = i;
        rank++;
        if (i <= n - 2 && part_id[i + 1] != part_id[i]) rank = 0;
    }
    int index = 0;
    for (int i = 0; i < M; ++i)
        if (part_map[i] == index) { real_map[index] = i; index++; }
    return index == n ? M : -2;
}

int or_mas_num_levels(int nn, int bank) {
    int nLevel = 1;
    int levelSz = (nn + bank - 1) / bank * bank;
    while (levelSz > bank) {
        levelSz /= bank;
        nLevel++;
        levelSz = (levelSz + bank - 1) / bank * bank;
    }
    nLevel = nLevel + 1;
    return nLevel > 6 ? 6 : nLevel;
}

/* ======================================================================== */
/* MAS: aggregation tables, Galerkin bank matrices, inverses, apply         */
/* ======================================================================== */

or_mas* or_mas_create(int n, int M, const int* part_map, const int* real_map,
                      const uint32_t* nstart, const uint32_t* nnum, const uint32_t* nlist,
                      int bank) {
    or_mas* m = (or_mas*)calloc(1, sizeof(or_mas));
    m->n = n; m->M = M; m->B = bank;
    m->L = or_mas_num_levels(M > n ? M : n, bank);
    m->part_map = (int*)malloc(sizeof(int) * (size_t)M);
    m->real_map = (int*)malloc(sizeof(int) * (size_t)n);
    memcpy(m->part_map, part_map, sizeof(int) * (size_t)M);
    memcpy(m->real_map, real_map, sizeof(int) * (size_t)n);
    int len = n > 0 ? (int)(nstart[n - 1] + nnum[n - 1]) : 0;
    m->nlist_len = len;
    m->nstart = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n + 1));
    m->nnum0 = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n + 1));
    m->nnum = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n + 1));
    m->nlist0 = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(len + 1));
    m->nlist = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(len + 1));
    memcpy(m->nstart, nstart, sizeof(uint32_t) * (size_t)n);
    memcpy(m->nnum0, nnum, sizeof(uint32_t) * (size_t)n);
    memcpy(m->nlist0, nlist, sizeof(uint32_t) * (size_t)len);
    m->level_size = (int*)calloc(2 * (size_t)(m->L + 2), sizeof(int));
    m->coarse_space = (int*)calloc((size_t)m->L * n + 1, sizeof(int));
    m->going_cap = 2 * M + m->L * n + 64;
    m->going_next = (int*)calloc((size_t)m->going_cap, sizeof(int));
    m->coarse_table = (int*)calloc(6 * (size_t)n + 1, sizeof(int));
    m->fine_mask = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
    m->prefix0 = (int*)calloc((size_t)M / bank + 1, sizeof(int));
    return m;
}

void or_mas_destroy(or_mas* m) {
    if (!m) return;
    free(m->part_map); free(m->real_map); free(m->nstart); free(m->nnum0); free(m->nnum);
    free(m->nlist0); free(m->nlist); free(m->level_size); free(m->coarse_space);
    free(m->going_next); free(m->coarse_table); free(m->fine_mask); free(m->prefix0);
    free(m->A); free(m->Ainv); free(m->AinvF);
    free(m->node_vptr); free(m->node_vlist);
    free(m);
}

/* decode an int4 contact pair into its vertex ids (MASPreconditioner.cu:1132-1540) */
static int decode_pair(const int* q, int node_offset, int v[4]) {
    int x = q[0], y = q[1], z = q[2], w = q[3], cnt;
    if (x >= 0) {
        if (w < 0) w = -w - 1;
        v[0] = x; v[1] = y; v[2] = z; v[3] = w; cnt = 4;
    } else {
        x = -x - 1;
        if (z < 0) {
            if (y < 0) { y = -y - 1; z = -z - 1; w = -w - 1; v[0] = x; v[1] = y; v[2] = z; v[3] = w; cnt = 4; }
            else { v[0] = x; v[1] = y; cnt = 2; }
        } else if (w < 0) {
            if (y < 0) { y = -y - 1; w = -w - 1; v[0] = x; v[1] = y; v[2] = z; v[3] = w; cnt = 4; }
            else { v[0] = x; v[1] = y; v[2] = z; cnt = 3; }
        } else { v[0] = x; v[1] = y; v[2] = z; v[3] = w; cnt = 4; }
    }
    for (int i = 0; i < cnt; ++i) v[i] -= node_offset;
    return cnt;
}

/* mask-OR collision connectivity into dst, ids either slots (level 0,
 * dst indexed by vertex) or level-local coarse ids (dst indexed by id). */
static void collision_connect(or_mas* m, uint32_t* dst, int level, const int* pairs, int np,
                              int node_offset) {
    int B = m->B, n = m->n;
    for (int p = 0; p < np; ++p) {
        int v[4], cp[4];
        int c = decode_pair(pairs + 4 * p, node_offset, v);
        for (int i = 0; i < c; ++i) {
            if (v[i] < 0 || v[i] >= n) cp[i] = -1;
            else if (level <= 0) cp[i] = m->real_map[v[i]];
            else cp[i] = m->coarse_space[(size_t)(level - 1) * n + v[i]];
        }
        uint32_t msk[4] = {0, 0, 0, 0};
        for (int i = 0; i < c; ++i)
            for (int j = i + 1; j < c; ++j) {
                if (cp[i] < 0 || cp[j] < 0 || cp[i] == cp[j]) continue;
                if (cp[i] / B == cp[j] / B) {
                    msk[i] |= 1u << (cp[j] % B);
                    msk[j] |= 1u << (cp[i] % B);
                }
            }
        for (int i = 0; i < c; ++i) {
            if (cp[i] < 0) continue;
            if (level <= 0) dst[v[i]] |= msk[i];
            else dst[cp[i]] |= msk[i];
        }
    }
}

static int popc(uint32_t x) { int c = 0; while (x) { x &= x - 1; ++c; } return c; }
static int ffs0(uint32_t x) { int i = 0; while (!(x & 1u)) { x >>= 1; ++i; } return i; }
static uint32_t lt_mask(int lane) { return lane >= 32 ? 0xffffffffu : ((1u << lane) - 1u); }

/* closure of `self` lane mask through masks[] of its bank (BFS over bits) */
static uint32_t closure(uint32_t own, int lane, const uint32_t* bank_masks) {
    uint32_t conn = own, visited = 1u << lane;
    for (;;) {
        uint32_t todo = visited ^ conn;
        if (!todo) break;
        int nv = ffs0(todo);
        visited |= 1u << nv;
        conn |= bank_masks[nv];
    }
    return conn;
}

static void reorder(or_mas* m, const int* pairs, int np, int node_offset) {
    int n = m->n, M = m->M, B = m->B, L = m->L;
    int nb0 = M / B;
    memcpy(m->nnum, m->nnum0, sizeof(uint32_t) * (size_t)n);
    memcpy(m->nlist, m->nlist0, sizeof(uint32_t) * (size_t)m->nlist_len);
    memset(m->level_size, 0, sizeof(int) * 2 * (size_t)(L + 2));
    /* level-0 connect masks (_buildCML0_new :63-106) */
    for (int s = 0; s < M; ++s) {
        int v = m->part_map[s];
        if (v < 0) continue;
        int bank = s / B, lane = s % B, nk = 0;
        uint32_t msk = 1u << lane, st = m->nstart[v];
        for (uint32_t i = 0; i < m->nnum[v]; ++i) {
            uint32_t u = m->nlist[st + i];
            int su = m->real_map[u];
            if (su / B == bank) msk |= 1u << (su % B);
            else m->nlist[st + nk++] = u;
        }
        m->nnum[v] = (uint32_t)nk;
        m->fine_mask[v] = msk;
    }
    if (np > 0) collision_connect(m, m->fine_mask, -1, pairs, np, node_offset);
    /* closure + component count per bank (_preparePrefixSumL0_new :157-214) */
    uint32_t* bm = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)B);
    uint32_t* cl = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)B);
    for (int b = 0; b < nb0; ++b) {
        for (int l = 0; l < B; ++l) {
            int v = m->part_map[b * B + l];
            bm[l] = v >= 0 ? m->fine_mask[v] : 0u;
        }
        int cnt = 0;
        for (int l = 0; l < B; ++l) {
            int v = m->part_map[b * B + l];
            if (v < 0) continue;
            cl[l] = closure(bm[l], l, bm);
            m->fine_mask[v] = cl[l];
            if (popc(cl[l] & lt_mask(l)) == 0) cnt++;
        }
        m->prefix0[b] = cnt;
    }
    /* level 1 numbering (_buildLevel1_new :271-328) */
    int* psum = (int*)malloc(sizeof(int) * ((size_t)nb0 + 1));
    int acc = 0;
    for (int b = 0; b < nb0; ++b) { psum[b] = acc; acc += m->prefix0[b]; }
    m->level_size[2 * 1 + 0] = acc;
    m->level_size[2 * 1 + 1] = (M + B - 1) / B * B;
    int Mpad = (M + B - 1) / B * B;
    for (int b = 0; b < nb0; ++b) {
        uint32_t elected = 0;
        for (int l = 0; l < B; ++l) {
            int v = m->part_map[b * B + l];
            if (v < 0) continue;
            if (popc(m->fine_mask[v] & lt_mask(l)) == 0) elected |= 1u << l;
        }
        for (int l = 0; l < B; ++l) {
            int v = m->part_map[b * B + l];
            if (v < 0) continue;
            int el = ffs0(m->fine_mask[v]);
            int pre = popc(elected & lt_mask(el)) + psum[b];
            m->coarse_space[v] = pre;
            m->going_next[v] = pre + Mpad;
        }
    }
    /* coarser levels */
    uint32_t* nmask = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
    int* npre = (int*)calloc((size_t)n + 1, sizeof(int));
    int* npsum = (int*)calloc((size_t)n + 1, sizeof(int));
    for (int level = 1; level < L; ++level) {
        memset(nmask, 0, sizeof(uint32_t) * (size_t)n);
        /* _buildConnectMaskLx_new :401-477 */
        for (int s = 0; s < M; ++s) {
            int v = m->part_map[s];
            if (v < 0) continue;
            int cidx = m->coarse_space[(size_t)(level - 1) * n + v], nk = 0;
            uint32_t msk = 0, st = m->nstart[v];
            for (uint32_t i = 0; i < m->nnum[v]; ++i) {
                uint32_t u = m->nlist[st + i];
                int cu = m->coarse_space[(size_t)(level - 1) * n + u];
                if (cidx / B == cu / B) msk |= 1u << (cu % B);
                else m->nlist[st + nk++] = u;
            }
            m->nnum[v] = (uint32_t)nk;
            nmask[cidx] |= msk;
        }
        if (np > 0) collision_connect(m, nmask, level, pairs, np, node_offset);
        int number = m->level_size[2 * level + 0];
        int begin = m->level_size[2 * level + 1];
        int nbk = (number + B - 1) / B;
        /* _nextLevelCluster :480-527 */
        for (int b = 0; b < nbk; ++b) {
            int cnt = 0;
            for (int l = 0; l < B; ++l) {
                int c = b * B + l;
                bm[l] = c < number ? ((1u << l) | nmask[c]) : 0u;
            }
            for (int l = 0; l < B; ++l) {
                int c = b * B + l;
                if (c >= number) continue;
                cl[l] = closure(bm[l], l, bm);
                if (popc(cl[l] & lt_mask(l)) == 0) cnt++;
            }
            for (int l = 0; l < B; ++l) {
                int c = b * B + l;
                if (c < number) nmask[c] = cl[l];
            }
            npre[b] = cnt;
        }
        /* _prefixSumLx :530-577 */
        acc = 0;
        for (int b = 0; b < nbk; ++b) { npsum[b] = acc; acc += npre[b]; }
        int pad = (number + B - 1) / B * B;
        m->level_size[2 * (level + 1) + 0] = acc;
        m->level_size[2 * (level + 1) + 1] = begin + pad;
        for (int b = 0; b < nbk; ++b) {
            uint32_t elected = 0;
            for (int l = 0; l < B; ++l) {
                int c = b * B + l;
                if (c < number && popc(nmask[c] & lt_mask(l)) == 0) elected |= 1u << l;
            }
            for (int l = 0; l < B; ++l) {
                int c = b * B + l;
                if (c >= number) continue;
                int el = ffs0(nmask[c]);
                int pre = popc(elected & lt_mask(el)) + npsum[b];
                bm[l] = (uint32_t)pre; /* temp */
            }
            for (int l = 0; l < B; ++l) {
                int c = b * B + l;
                if (c >= number) continue;
                nmask[c] = bm[l];
                if (begin + c < m->going_cap) m->going_next[begin + c] = (int)bm[l] + begin + pad;
            }
        }
        /* _computeNextLevel :579-589 */
        for (int v = 0; v < n; ++v) {
            int nx = m->coarse_space[(size_t)(level - 1) * n + v];
            m->coarse_space[(size_t)level * n + v] = (int)nmask[nx];
        }
    }
    m->total = m->level_size[2 * L + 1];
    /* _aggregationKernel :592-627 */
    for (int v = 0; v < n; ++v) {
        int cur = v;
        for (int l = 0; l < L - 1; ++l) {
            int nx = m->going_next[cur];
            cur = nx;
            m->coarse_table[6 * (size_t)v + l] = nx;
        }
    }
    free(bm); free(cl); free(psum); free(nmask); free(npre); free(npsum);
}

/* coarse node (global id) of vertex v at level l (l=0: slot) */
int or__node_of(const or_mas* m, int v, int l) {
    return l == 0 ? m->real_map[v] : m->coarse_table[6 * (size_t)v + l - 1];
}
static int node_of(const or_mas* m, int v, int l) { return or__node_of(m, v, l); }

/* unpivoted in-place Gauss-Jordan inverse of bank b (:630-726); S: (3B)^2
 * scratch, colm: 3B scratch */
void or__invert_bank(or_mas* m, int b, double* S, double* colm) {
    int D = 3 * m->B;
    size_t bs = (size_t)D * D;
    memcpy(S, m->A + bs * b, sizeof(double) * bs);
    for (int i = 0; i < D; ++i) if (S[(size_t)i * D + i] == 0.0) S[(size_t)i * D + i] = 1.0;
    for (int j = 0; j < D; ++j) {
        double rt = S[(size_t)j * D + j];
        for (int i = 0; i < D; ++i) colm[i] = S[(size_t)i * D + j];
        for (int i = 0; i < D; ++i) S[(size_t)i * D + j] = (i == j) ? 1.0 : 0.0;
        for (int i = 0; i < D; ++i) S[(size_t)j * D + i] /= rt;
        for (int k = 0; k < D; ++k) {
            if (k == j) continue;
            double rate = -colm[k];
            for (int i = 0; i < D; ++i) S[(size_t)k * D + i] += rate * S[(size_t)j * D + i];
        }
    }
    double* Ai = m->Ainv + bs * b;
    float* Af = m->AinvF + bs * b;
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            double v = i <= j ? S[(size_t)i * D + j] : S[(size_t)j * D + i];
            Ai[(size_t)i * D + j] = v;
            Af[(size_t)i * D + j] = (float)v;
        }
}

int or_mas_build(or_mas* m, int K, const int* rows, const int* cols, const double* vals,
                 int offset, const int* pairs, int npairs, int node_offset, int fp32_store) {
    if (m->n < 1) return 0;
    reorder(m, pairs, npairs, node_offset);
    int B = m->B, D = 3 * B, nbanks = m->total / B;
    size_t bs = (size_t)D * D;
    if (nbanks > m->nbanks_alloc) {
        free(m->A); free(m->Ainv); free(m->AinvF);
        m->A = (double*)malloc(sizeof(double) * bs * nbanks);
        m->Ainv = (double*)malloc(sizeof(double) * bs * nbanks);
        m->AinvF = (float*)malloc(sizeof(float) * bs * nbanks);
        m->nbanks_alloc = nbanks;
    }
    memset(m->A, 0, sizeof(double) * bs * nbanks);
    m->fp32_store = fp32_store;
    m->node_lists_ok = 0;
    if (or_threads > 1) {
        or_mas_project_invert_mt(m, K, rows, cols, vals, offset);
        return m->total;
    }
    /* Galerkin projection A_l = R_l A R_l^T restricted to same-bank blocks
     * (PrepareHessian_bcoo, MASPreconditioner.cu:1816-2062) */
    for (int l = 0; l < m->L; ++l) {
        for (int k = 0; k < K; ++k) {
            int r = rows[k] - offset, c = cols[k] - offset;
            if (r < 0 || c < 0 || r >= m->n || c >= m->n) continue;
            const double* H = vals + 9 * (size_t)k; /* col-major */
            for (int tr = 0; tr < 2; ++tr) {
                if (tr == 1 && r == c) break;
                int I = node_of(m, tr ? c : r, l), J = node_of(m, tr ? r : c, l);
                if (I / B != J / B) continue;
                double* A = m->A + bs * (size_t)(I / B);
                int bi = (I % B) * 3, bj = (J % B) * 3;
                for (int p = 0; p < 3; ++p)
                    for (int q = 0; q < 3; ++q) {
                        double h = tr ? H[q + 3 * p] : H[p + 3 * q];
                        A[(size_t)(bi + p) * D + bj + q] += h;
                    }
            }
        }
    }
    /* unpivoted in-place Gauss-Jordan inverse per bank (:630-726) */
    double* S = (double*)malloc(sizeof(double) * bs);
    double* colm = (double*)malloc(sizeof(double) * D);
    for (int b = 0; b < nbanks; ++b) or__invert_bank(m, b, S, colm);
    free(S); free(colm);
    return m->total;
}

void or_mas_apply(or_mas* m, const double* r, double* z, int fp32) {
    int n = m->n, B = m->B, D = 3 * B, L = m->L, tot = m->total;
    size_t bs = (size_t)D * D;
    if (n < 1) return;
    if (!fp32) {
        double* R = (double*)calloc(3 * (size_t)tot, sizeof(double));
        double* Z = (double*)calloc(3 * (size_t)tot, sizeof(double));
        for (int l = 0; l < L; ++l)
            for (int v = 0; v < n; ++v) {
                int I = node_of(m, v, l);
                for (int k = 0; k < 3; ++k) R[3 * (size_t)I + k] += r[3 * (size_t)v + k];
            }
        for (int b = 0; b < tot / B; ++b) {
            const double* Ai = m->Ainv + bs * b;
            for (int i = 0; i < D; ++i) {
                double s = 0;
                for (int j = 0; j < D; ++j) s += Ai[(size_t)i * D + j] * R[(size_t)b * D + j];
                Z[(size_t)b * D + i] = s;
            }
        }
        for (int v = 0; v < n; ++v)
            for (int k = 0; k < 3; ++k) {
                double s = 0;
                for (int l = 0; l < L; ++l) s += Z[3 * (size_t)node_of(m, v, l) + k];
                z[3 * (size_t)v + k] = s;
            }
        free(R); free(Z);
    } else {
        float* R = (float*)calloc(3 * (size_t)tot, sizeof(float));
        float* Z = (float*)calloc(3 * (size_t)tot, sizeof(float));
        for (int l = 0; l < L; ++l)
            for (int v = 0; v < n; ++v) {
                int I = node_of(m, v, l);
                for (int k = 0; k < 3; ++k) R[3 * (size_t)I + k] += (float)r[3 * (size_t)v + k];
            }
        for (int b = 0; b < tot / B; ++b) {
            const float* Ai = m->AinvF + bs * b;
            for (int i = 0; i < D; ++i) {
                float s = 0;
                for (int j = 0; j < D; ++j) s += Ai[(size_t)i * D + j] * R[(size_t)b * D + j];
                Z[(size_t)b * D + i] = s;
            }
        }
        for (int v = 0; v < n; ++v)
            for (int k = 0; k < 3; ++k) {
                float s = 0;
                for (int l = 0; l < L; ++l) s += Z[3 * (size_t)node_of(m, v, l) + k];
                z[3 * (size_t)v + k] = (double)s;
            }
        free(R); free(Z);
    }
}

int or_mas_levelnum(const or_mas* m) { return m->L; }
int or_mas_total_clusters(const or_mas* m) { return m->total; }

void or_mas_tables(const or_mas* m, int* level_size, int* coarse_space, int* going_next,
                   int* coarse_table, uint32_t* fine_mask, int* prefix0) {
    int n = m->n, L = m->L;
    if (level_size) memcpy(level_size, m->level_size, sizeof(int) * 2 * (size_t)(L + 1));
    if (coarse_space) memcpy(coarse_space, m->coarse_space, sizeof(int) * (size_t)L * n);
    if (going_next) {
        /* goingNext entries: [0,n) for vertices, [begin_l, begin_l+count_l) per level */
        memcpy(going_next, m->going_next, sizeof(int) * (size_t)(m->total > n ? m->total : n));
    }
    if (coarse_table) memcpy(coarse_table, m->coarse_table, sizeof(int) * 6 * (size_t)n);
    if (fine_mask) memcpy(fine_mask, m->fine_mask, sizeof(uint32_t) * (size_t)n);
    if (prefix0) memcpy(prefix0, m->prefix0, sizeof(int) * (size_t)(m->M / m->B));
}

void or_mas_bank_matrix(const or_mas* m, int b, double* A, double* Ainv) {
    size_t bs = (size_t)(3 * m->B) * (3 * m->B);
    if (A) memcpy(A, m->A + bs * b, sizeof(double) * bs);
    if (Ainv) memcpy(Ainv, m->Ainv + bs * b, sizeof(double) * bs);
}

/* ======================================================================== */
/* PCG                                                                      */
/* ======================================================================== */
static double dot(const double* a, const double* b, int n) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static void inv3(const double* M /*col-major*/, double* R /*col-major*/) {
    /* element (r,c) = M[r + 3c] */
    double a = M[0], b = M[3], c = M[6], d = M[1], e = M[4], f = M[7], g = M[2], h = M[5], i = M[8];
    double A = e * i - f * h, Bq = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * Bq + c * C, id = 1.0 / det;
    double r00 = A * id, r01 = -(b * i - c * h) * id, r02 = (b * f - c * e) * id;
    double r10 = Bq * id, r11 = (a * i - c * g) * id, r12 = -(a * f - c * d) * id;
    double r20 = C * id, r21 = -(a * h - b * g) * id, r22 = (a * e - b * d) * id;
    R[0] = r00; R[3] = r01; R[6] = r02; R[1] = r10; R[4] = r11; R[7] = r12; R[2] = r20; R[5] = r21; R[8] = r22;
}

static void apply_pre(int precond, int nb, const double* Dinv, or_mas* mas, int mas_fp32,
                      const double* r, double* z) {
    if (precond == 2 && mas) { or_mas_apply(mas, r, z, mas_fp32); return; }
    if (precond == 1) {
        for (int i = 0; i < nb; ++i)
            for (int p = 0; p < 3; ++p) {
                double s = 0;
                for (int q = 0; q < 3; ++q) s += Dinv[9 * (size_t)i + p + 3 * q] * r[3 * i + q];
                z[3 * i + p] = s;
            }
        return;
    }
    memcpy(z, r, sizeof(double) * 3 * (size_t)nb);
}

int or_pcg(int K, const int* rows, const int* cols, const double* vals, int nb, const double* b,
           double* x, int precond, or_mas* mas, int mas_fp32, double tol, int rule, int max_iter,
           double* rel_res_out) {
    if (or_threads > 1)
        return or_pcg_mt(K, rows, cols, vals, nb, b, x, precond, mas, mas_fp32, tol, rule, max_iter, rel_res_out);
    int N = 3 * nb;
    double *r = (double*)malloc(sizeof(double) * N), *z = (double*)malloc(sizeof(double) * N);
    double *p = (double*)malloc(sizeof(double) * N), *Ap = (double*)malloc(sizeof(double) * N);
    double* Dinv = NULL;
    if (precond == 1) {
        Dinv = (double*)calloc(9 * (size_t)nb, sizeof(double));
        for (int k = 0; k < K; ++k)
            if (rows[k] == cols[k]) inv3(vals + 9 * (size_t)k, Dinv + 9 * (size_t)rows[k]);
    }
    memset(x, 0, sizeof(double) * N);
    memcpy(r, b, sizeof(double) * N);
    double bn = sqrt(dot(b, b, N));
    apply_pre(precond, nb, Dinv, mas, mas_fp32, r, z);
    double rz = dot(r, z, N), rz0 = rz;
    memcpy(p, z, sizeof(double) * N);
    int k;
    for (k = 1; k < max_iter; ++k) {
        or_spmv_sym(K, rows, cols, vals, nb, 1.0, p, 0.0, Ap);
        double pAp = dot(p, Ap, N), alpha = rz / pAp;
        for (int i = 0; i < N; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
        if (rule == 0) {
            if (fabs(rz) <= tol * rz0) break;
        } else {
            if (sqrt(dot(r, r, N)) <= tol * bn) break;
        }
        apply_pre(precond, nb, Dinv, mas, mas_fp32, r, z);
        double rzn = dot(r, z, N), beta = rzn / rz;
        for (int i = 0; i < N; ++i) p[i] = z[i] + beta * p[i];
        rz = rzn;
    }
    if (rel_res_out) {
        or_spmv_sym(K, rows, cols, vals, nb, 1.0, x, 0.0, Ap);
        double s = 0;
        for (int i = 0; i < N; ++i) { double d = b[i] - Ap[i]; s += d * d; }
        *rel_res_out = bn > 0 ? sqrt(s) / bn : sqrt(s);
    }
    free(r); free(z); free(p); free(Ap); free(Dinv);
    return k;
}
