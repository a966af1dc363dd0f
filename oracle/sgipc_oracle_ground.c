/*
 * sgipc_oracle_ground.c -- CPU restatement (FP64, C99) of the ground-plane
 * barrier and of the SNK1 parameter adjoint.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the library's ground and
 * adjoint kernels; nothing in the product links or calls it.
 *
 * Ground barrier, plane n.x = o, s = (n.x - o)^2, dhat squared:
 *   energy kappa * sum_{s < dhat} -(s - dhat)^2 ln(s / dhat)
 *       (StiffGIPC/GIPC.cu:6837-6900, kappa after the sum :10555)
 *   gradient kappa g_b 2 d n, Hessian kappa (4 H_b s + 2 g_b) n n^T
 *       (StiffGIPC/GIPC.cu:6463-6514)
 *   detection s <= dhat (StiffGIPC/GIPC.cu:6433-6450), kept in list order
 *   step bound 1 / max(1, max_{n.dir > 0} n.dir / (slack d))
 *       (StiffGIPC/GIPC.cu:6902-6960, :8986-9030)
 * Parameter adjoint: d(adj . g)/d mu and d(adj . g)/d lam of the SNK1 tet
 * gradient g = dt^2 vol Bm P^T, P = mu F + (lam (J - 1) - mu) cof(F)
 * (StiffGIPC/femEnergy.cu:432-459, :1719-1857); no reference counterpart.
 */
#include "sgipc_oracle.h"
#include <math.h>

static double pdist(const double* x, uint32_t v, const double n[3], double o) {
    return n[0] * x[3 * v] + n[1] * x[3 * v + 1] + n[2] * x[3 * v + 2] - o;
}

int or_ground_detect(int m, const uint32_t* cand, const double* x, const double nrm[3], double o, double dhat,
                     uint32_t* out) {
    int c = 0;
    for (int i = 0; i < m; ++i) {
        double d = pdist(x, cand[i], nrm, o);
        if (d * d <= dhat) out[c++] = cand[i];
    }
    return c;
}

void or_ground_assemble(int np, const uint32_t* gp, const double* x, const double nrm[3], double o, double dhat,
                        double kappa, const int* btype, int off, double* fb, double* vals, int* rows, int* cols) {
    for (int i = 0; i < np; ++i) {
        uint32_t v = gp[i];
        double d = pdist(x, v, nrm, o), s = d * d, t = s - dhat, lg = log(s / dhat);
        double gb = -2.0 * t * lg - t * t / s;
        double hb = -2.0 * lg - 4.0 * t / s + t * t / (s * s);
        int pin = btype ? btype[v] != 0 : 0;
        double gs = pin ? 0.0 : kappa * gb * 2.0 * d;
        double hs = pin ? 0.0 : kappa * (4.0 * hb * s + 2.0 * gb);
        for (int k = 0; k < 3; ++k) fb[3 * v + k] += gs * nrm[k];
        for (int q = 0; q < 3; ++q)
            for (int p = 0; p < 3; ++p) vals[9 * i + p + 3 * q] = hs * nrm[p] * nrm[q];
        rows[i] = (int)v + off;
        cols[i] = (int)v + off;
    }
}

double or_ground_energy(int m, const uint32_t* ids, const double* x, const double nrm[3], double o, double dhat,
                        double kappa) {
    double e = 0.0;
    for (int i = 0; i < m; ++i) {
        double d = pdist(x, ids[i], nrm, o), s = d * d;
        if (s < dhat) e += -(s - dhat) * (s - dhat) * log(s / dhat);
    }
    return kappa * e;
}

double or_ground_step_bound(int m, const uint32_t* ids, const double* x, const double* dir, const double nrm[3],
                            double o, double slack) {
    double r = 1.0;
    for (int i = 0; i < m; ++i) {
        uint32_t v = ids[i];
        double coef = nrm[0] * dir[3 * v] + nrm[1] * dir[3 * v + 1] + nrm[2] * dir[3 * v + 2];
        if (coef > 0.0) {
            double q = coef / (pdist(x, v, nrm, o) * slack);
            if (q > r) r = q;
        }
    }
    return 1.0 / r;
}

void or_fem_param_adjoint(int T, const double* x, const uint32_t* tets, const double* DmInv, const double* vol,
                          double dt, const int* btype, const double* adj, double* dmu, double* dlam) {
    for (int t = 0; t < T; ++t) {
        const uint32_t* v = tets + 4 * t;
        const double* D = DmInv + 9 * t;
        double Ds[9], F[9], Bm[12], A[12], G[9], cof[9];
        for (int i = 0; i < 3; ++i)
            for (int a = 0; a < 3; ++a) Ds[3 * i + a] = x[3 * v[a + 1] + i] - x[3 * v[0] + i];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                F[3 * i + j] = Ds[3 * i] * D[j] + Ds[3 * i + 1] * D[3 + j] + Ds[3 * i + 2] * D[6 + j];
        for (int j = 0; j < 3; ++j) {
            Bm[j] = -(D[j] + D[3 + j] + D[6 + j]);
            for (int a = 0; a < 3; ++a) Bm[3 * (a + 1) + j] = D[3 * a + j];
        }
        for (int a = 0; a < 4; ++a) {
            int pin = btype ? btype[v[a]] != 0 : 0;
            for (int k = 0; k < 3; ++k) A[3 * a + k] = pin ? 0.0 : adj[3 * v[a] + k];
        }
        /* G(k, j) = sum_a adj_a(k) Bm(a, j), so adj . g = dt^2 vol <G, P> */
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) {
                double g = 0.0;
                for (int a = 0; a < 4; ++a) g += A[3 * a + k] * Bm[3 * a + j];
                G[3 * k + j] = g;
            }
        cof[0] = F[4] * F[8] - F[5] * F[7]; cof[1] = F[5] * F[6] - F[3] * F[8]; cof[2] = F[3] * F[7] - F[4] * F[6];
        cof[3] = F[7] * F[2] - F[8] * F[1]; cof[4] = F[8] * F[0] - F[6] * F[2]; cof[5] = F[6] * F[1] - F[7] * F[0];
        cof[6] = F[1] * F[5] - F[4] * F[2]; cof[7] = F[2] * F[3] - F[0] * F[5]; cof[8] = F[0] * F[4] - F[1] * F[3];
        double J = F[0] * cof[0] + F[1] * cof[1] + F[2] * cof[2];
        double gf = 0.0, gc = 0.0;
        for (int i = 0; i < 9; ++i) { gf += G[i] * F[i]; gc += G[i] * cof[i]; }
        double s = dt * dt * vol[t];
        dmu[t] = s * (gf - gc);
        dlam[t] = s * (J - 1.0) * gc;
    }
}
