/*
 * sgipc_oracle.h -- CPU restatement (FP64, plain C) of Stiff-GIPC's Newton-step
 * hot path.  TEST INFRASTRUCTURE ONLY: this library is the checker that the HIP
 * product path is compared against, and the CPU baseline timed beside it.  The
 * product (libsgipc) never links, loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root, S/ = StiffGIPC/).
 *
 * Conventions (same as the reference):
 *   - vertices: double3 AoS, x[3*i+k]
 *   - tets: uint32 x4 (uint4)
 *   - DmInv: 3x3 row-major per tet (Matrix3x3d m[3][3])
 *   - triplet / BCOO block values: 9 doubles, column-major (Eigen::Matrix3d)
 *   - BCOO: upper block triangle (row <= col), sorted by (row, col)
 */
#ifndef SGIPC_ORACLE_H
#define SGIPC_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- small dense math --------------------------------------------------- */
/* Signed 3x3 SVD A = U diag(S) V^T, U,V proper rotations (implicit-shift QR,
 * S/QRSVD.hpp:219-317, S/givens.hpp).  A,U,V row-major. */
void or_svd3(const double A[9], double U[9], double S[3], double V[9]);
/* Symmetric 3x3 eigen-decomposition, ascending eigenvalues, eigenvectors in
 * columns of Q (row-major).  Stands in for Eigen::SelfAdjointEigenSolver::
 * computeDirect used at S/femEnergy.cu:1697-1698. */
void or_syev3(const double A[9], double w[3], double Q[9]);

/* ---- SNK1 tet (S/femEnergy.cu:1032-1050, 432-459, 1389-1436, 1593-1857) --- */
double or_snk1_energy(const double xt[12], const double DmInv[9], double vol,
                      double mu, double lam);
/* grad12 = dt^2 vol PFPX^T vec(P); H144 = dt^2 vol PFPX^T Hq+ PFPX (row-major 12x12) */
void or_snk1_tet(const double xt[12], const double DmInv[9], double vol,
                 double mu, double lam, double dt, double grad12[12],
                 double H144[144]);
/* Projected 9x9 Hessian dPsi/dF (vec = column-major), H81 row-major. */
void or_snk1_hq(const double F[9], double mu, double lam, double H81[81]);

/* Rest-shape preprocessing (S/gl_main.cu:557-617). DmInv row-major. */
void or_fem_precompute(int n, int T, const double* xrest, const uint32_t* tets,
                       double E, double nu, double density, double* DmInv,
                       double* vol, double* mass, double* mu, double* lam);

/* FEM triplets + shape gradient (S/femEnergy.cu:1719-1857 + boundary mask
 * S/GIPC.cu:10295-10313).  grad (3n) is accumulated (+=) in tet order.
 * Writes 10 triplets per tet at vals+90*t, rows/cols+10*t. */
double or_accd(int pt, const double* v, const double* u, double eta);
double or_ccd_step_bound(int npt, const int* pt, int nee, const int* ee, const double* x, const double* dir,
                         double slack);
extern int or_fem_project;
void or_set_fem_projection(int on);
/* Threads of the CPU baseline: t > 1 runs or_fem_assemble, or_convert,
 * or_mas_build and or_pcg on t threads (sgipc_oracle_mt.c); 1 (default) the
 * serial code. */
void or_set_threads(int t);
int or_get_threads(void);
/* ground plane n.x = o (sgipc_oracle_ground.c) */
int or_ground_detect(int m, const uint32_t* cand, const double* x, const double nrm[3], double o, double dhat,
                     uint32_t* out);
void or_ground_assemble(int np, const uint32_t* gp, const double* x, const double nrm[3], double o, double dhat,
                        double kappa, const int* btype, int off, double* fb, double* vals, int* rows, int* cols);
double or_ground_energy(int m, const uint32_t* ids, const double* x, const double nrm[3], double o, double dhat,
                        double kappa);
double or_ground_step_bound(int m, const uint32_t* ids, const double* x, const double* dir, const double nrm[3],
                            double o, double slack);
void or_fem_param_adjoint(int T, const double* x, const uint32_t* tets, const double* DmInv, const double* vol,
                          double dt, const int* btype, const double* adj, double* dmu, double* dlam);
void or_fem_assemble(int n, int T, const double* x, const uint32_t* tets,
                     const double* DmInv, const double* vol, const double* mu,
                     const double* lam, double dt, const int* btype,
                     int index_offset, double* grad, double* vals, int* rows,
                     int* cols);

/* Kinetic gradient g = m (x - xtilde) (S/GIPC.cu:6339-6350). */
void or_kinetic_gradient(int n, const double* x, const double* xt,
                         const double* mass, double* g);
/* Mass diagonal triplets (S/GIPC.cu:10317-10336). */
void or_mass_triplets(int n, const double* mass, int index_offset, double* vals,
                      int* rows, int* cols);
/* RHS b = fb + shape (0 on boundary vertices) (fem_linear_subsystem.cu:66-83) */
void or_fem_rhs(int n, const double* fb, const double* shape, const int* btype,
                double* b);

/* ---- triplet -> BCOO (S/linear_system/utils/converter.cu:23-196) ---------
 * Returns K; outputs sorted by (row,col); each block summed in ascending
 * triplet-index order. out arrays must hold nt entries. */
int or_convert(int nt, const int* rows, const int* cols, const double* vals,
               int* out_rows, int* out_cols, double* out_vals);

/* y = a*A*x + b*y, A = U + U^T - diag(U), upper BCOO (spmv.cu:8-133). */
void or_spmv_sym(int K, const int* rows, const int* cols, const double* vals,
                 int nb, double a, const double* x, double b, double* y);

/* ---- MAS neighbour / bank inputs ---------------------------------------- */
/* Vertex neighbour CSR from tets (S/load_mesh.cpp:934-1040). start/num are n
 * entries, list must hold 12*T entries.  Returns list length. */
int or_vertex_neighbors(int n, int T, const uint32_t* tets, uint32_t* start,
                        uint32_t* num, uint32_t* list);
/* Bank map from ascending part ids (S/gl_main.cu:1066-1093).  Returns
 * map size M = 16*nparts (or bank*nparts); part_map must hold that many. */
int or_bank_map(int n, const int* part_id, int bank, int* part_map,
                int* real_map);
int or_mas_num_levels(int n_nodes, int bank); /* MASPreconditioner.cu:1724-1743 */

typedef struct or_mas or_mas;
/* bank must be 16 or 32. */
or_mas* or_mas_create(int n, int M, const int* part_map, const int* real_map,
                      const uint32_t* nstart, const uint32_t* nnum,
                      const uint32_t* nlist, int bank);
void or_mas_destroy(or_mas* m);
/* Aggregation tables (ReorderRealtime, MASPreconditioner.cu:1772-1813) +
 * Galerkin projection (PrepareHessian_bcoo :1816-2062) + Gauss-Jordan bank
 * inverses (__inverse6_P96x96 :630-726).  pairs: int4 contact pairs in the
 * reference sign encoding, node_offset subtracted from their vertex ids.
 * K blocks of upper BCOO restricted to the subsystem range, block indices
 * shifted by `offset`.  fp32_store: store inverses as float (reference). */
int or_mas_build(or_mas* m, int K, const int* rows, const int* cols,
                 const double* vals, int offset, const int* pairs, int npairs,
                 int node_offset, int fp32_store);
/* z = MAS(r).  fp32: reference FP32 r/z arithmetic, else FP64. */
void or_mas_apply(or_mas* m, const double* r, double* z, int fp32);
/* table accessors (bit-exact parity) */
int or_mas_levelnum(const or_mas* m);
int or_mas_total_clusters(const or_mas* m);
void or_mas_tables(const or_mas* m, int* level_size /*2*(L+1)*/,
                   int* coarse_space /*L*n*/, int* going_next /*L*n*/,
                   int* coarse_table /*6*n*/, uint32_t* fine_mask /*n*/,
                   int* prefix0 /*M/bank*/);
/* Dense FP64 Galerkin bank matrix (before inversion) and inverse (symmetrised
 * as used by the apply), 3b x 3b row-major, bank index over all levels. */
void or_mas_bank_matrix(const or_mas* m, int bank_index, double* A, double* Ainv);

/* ---- PCG (S/linear_system/solver/pcg_solver.cu:158-267) -----------------
 * precond: 0 none, 1 block-Jacobi 3x3 (diag_preconditioner.cu), 2 MAS.
 * rule 0: reference |rz| <= tol*rz0 (stale rz); rule 1: ||r||/||b|| <= tol.
 * max_iter: reference uses 0.3*dof.  Returns iterations (reference count). */
int or_pcg(int K, const int* rows, const int* cols, const double* vals, int nb,
           const double* b, double* x, int precond, or_mas* mas, int mas_fp32,
           double tol, int rule, int max_iter, double* rel_res_out);

/* ---- IPC barrier (S/GIPC.cu:3106-4943, energy :701-957) ------------------
 * Pair p is an int4 in the reference sign encoding.  dhat is the squared
 * barrier distance, kappa the stiffness.  Per pair: gradient (12 doubles,
 * vertex order of the decoded pair) and dense Hessian (row-major nv3 x nv3),
 * returns the number of vertices (2, 3 or 4), their ids in vid, the size
 * class (2: 6x6, 3: 9x9, 4: 12x12) in *cls, and 0 gradient/Hessian when a
 * mollified pair has zero edge cross product. */
int or_barrier_pair(const double* x, const double* xrest, const int q[4], double dhat, double kappa,
                    int vid[4], int* cls, double grad[12], double H[144], double* energy);
/* Assemble all pairs: fb += gradient, triplets at offset: class-4 pairs at
 * 16*matIndex, class-3 at 16*c4 + 9*matIndex, class-2 at 16*c4 + 9*c3 +
 * 4*matIndex (write_triplet, S/GIPC.cu:80-137), then the fem-fem contact
 * remap (row<=col kept, zeroed on pinned vertices; lower blocks swapped and
 * zeroed, S/GIPC.cu:10183-10214). counts[3] = {c2, c3, c4}. */
void or_barrier_assemble(int npairs, const int* pairs, const int* mat_index, const int counts[3],
                         const double* x, const double* xrest, double dhat, double kappa,
                         const int* btype, double* fb, double* vals, int* rows, int* cols);
/* Total barrier energy (sum over pairs, S/GIPC.cu:701-957). */
double or_barrier_energy(int npairs, const int* pairs, const double* x, const double* xrest,
                         double dhat, double kappa);
/* Narrow-phase classification and int4 encoding (S/mlbvh.cu:288-1060).
 * pt candidates (p, t0, t1, t2) and ee candidates (a0, a1, b0, b1); output
 * pairs of candidates with d < dhat, PT candidates first, each in candidate
 * order; mat_index = rank within the size class; counts = {c2, c3, c4}.
 * Returns the number of pairs. */
int or_classify(int npt, const int* pt, int nee, const int* ee, const double* x, const double* xrest,
                double dhat, int* pairs, int* mat_index, int counts[3]);

#ifdef __cplusplus
}
#endif
#endif
