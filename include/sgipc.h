/*
 * sgipc.h -- C ABI of the MI355X (gfx950) Newton-step library: FEM + barrier
 * assembly into block-3x3 triplets, triplet -> BCOO conversion, symmetric
 * SpMV, MAS / block-Jacobi preconditioners and PCG.
 *
 * It replaces the C++ linear-system boundary of Stiff-GIPC
 * (namespace gipc, StiffGIPC/linear_system/), keeping its data layout so
 * reference-format triplets can be fed unchanged:
 *   - block values: 9 doubles per 3x3 block, COLUMN-major (Eigen::Matrix3d);
 *   - block indices: int32 in 3-DOF units, upper triangle (row <= col);
 *   - vertices: double3 AoS; tets: uint32 x4; DmInv: 3x3 row-major per tet;
 *   - contact pairs: int4 in the reference sign encoding.
 *
 * Conventions: every function returns SGIPC_OK (0) or an error code and never
 * aborts; sgipc_last_error() describes the last failure on a context.  Device
 * pointers ("d_") are caller-owned device memory; host pointers ("h_") are
 * host memory.  Work is enqueued on the context's stream; functions that
 * return host values (counts, solver statistics) synchronise that stream.
 * A context is driven by one host thread at a time.
 */
#ifndef SGIPC_H
#define SGIPC_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SGIPC_OK 0
#define SGIPC_ERR_INVALID_ARG 1
#define SGIPC_ERR_HIP 2
#define SGIPC_ERR_CAPACITY 3
#define SGIPC_ERR_STATE 4
#define SGIPC_ERR_INTERNAL 5

#define SGIPC_PRECOND_NONE 0
#define SGIPC_PRECOND_DIAG 1     /* 3x3 block Jacobi (diag_preconditioner.cu) */
#define SGIPC_PRECOND_MAS 2      /* multilevel additive Schwarz, FP64 apply */
#define SGIPC_PRECOND_MAS_FP32 3 /* MAS with FP32 inverses/apply (reference precision) */

#define SGIPC_PRECISION_FP64 0
#define SGIPC_PRECISION_MIXED 1

#define SGIPC_STOP_REF_RZ 0      /* |r.z| <= tol * r0.z0 on the previous rz (pcg_solver.cu:232) */
#define SGIPC_STOP_TRUE_REL 1    /* ||r||_2 <= tol * ||b||_2 */

typedef struct sgipc_ctx sgipc_ctx;

typedef struct sgipc_solve_opts {
    int precision;        /* SGIPC_PRECISION_* */
    int stop_rule;        /* SGIPC_STOP_* */
    double tol;           /* reference default 1e-4 (parameterSetting.txt:14) */
    double max_iter_ratio;/* max iterations = ratio * dof (pcg_solver.h:13, default 0.3) */
    int max_iter;         /* >0 overrides max_iter_ratio */
    int check_every;      /* host polls the device stop flag every k iterations */
} sgipc_solve_opts;

typedef struct sgipc_solve_stats {
    int iterations;       /* reference iteration count (pcg_solver.cu:267) */
    int converged;
    int refinements;      /* outer refinement steps (mixed precision) */
    double rel_residual;  /* true ||b - A x|| / ||b||, recomputed in FP64 */
    double rz0;
} sgipc_solve_stats;

/* The sixteen values of the reference's parameterSetting.txt, in file order
 * (LoadSettings, StiffGIPC/gl_main.cu:677-726). */
typedef struct sgipc_parameters {
    double density, poisson_rate, friction_rate, gd_friction_rate;
    double cloth_thickness, cloth_youngs_modulus, bend_youngs_modulus, cloth_density;
    double strain_rate, soft_motion_rate, collision_detection_buff_scale, motion_rate;
    double dt, pcg_threshold, newton_threshold, relative_dhat;
} sgipc_parameters;

/* ---- context -------------------------------------------------------------- */
int sgipc_version(void);
/* replaces GlobalLinearSystem construction (global_linear_system.h:24) */
int sgipc_ctx_create(int device, sgipc_ctx** out);
int sgipc_ctx_destroy(sgipc_ctx* ctx);
/* run on a caller stream (hipStream_t); NULL restores the context's own, a
 * blocking stream (ordered with the legacy null stream) */
int sgipc_ctx_set_stream(sgipc_ctx* ctx, void* hip_stream);
int sgipc_ctx_synchronize(sgipc_ctx* ctx);
const char* sgipc_last_error(const sgipc_ctx* ctx);

/* ---- host-side topology preprocessing ------------------------------------- */
/* Vertex neighbour CSR from tets; replaces tetrahedra_obj::getVertNeighbors
 * (StiffGIPC/load_mesh.cpp:934-1040).  list must hold 12*T entries. */
int sgipc_build_vertex_neighbors(int n, int T, const uint32_t* h_tets, uint32_t* h_start,
                                 uint32_t* h_num, uint32_t* h_list, int list_capacity,
                                 int* list_len);
/* Bank slot maps from ascending part ids; replaces setMAS_partition
 * (StiffGIPC/gl_main.cu:1066-1093).  part_map needs bank*(last part+1). */
int sgipc_build_bank_map(int n, const int* h_part_id, int bank, int* h_part_map,
                         int part_map_capacity, int* map_size, int* h_real_map);
/* Vertex partition into subdomains of at most `bank` (16 or 32) vertices;
 * replaces metis_sort (MeshProcess/metis_partition/src/metis_sort.cpp:25-383:
 * nParts = ceil(n / (bank - off)), off raised until no part exceeds bank) with
 * this library's own multilevel recursive bisection of the vertex graph
 * (edge weight = tets sharing the edge, mesh.cpp:486-549) in place of
 * METIS_PartGraphKway (node_edge_model.cpp:5-28).  h_part[v] = part of vertex
 * v; h_order (optional, n entries) = vertex ids sorted by part, ascending id
 * within a part (TetMesh::sort_index, mesh.cpp:65-88, which sorts unstably):
 * new vertex i is old vertex h_order[i].  offset (optional) = the off at
 * which the loop stopped (it starts at 1: one vertex of slack per part). */
int sgipc_partition_mesh(int n, int T, const uint32_t* h_tets, int bank, int* h_part, int* h_order, int* n_parts,
                         int* offset);
/* Sum of the weights of the vertex-graph edges cut by a partition. */
int sgipc_partition_edge_cut(int n, int T, const uint32_t* h_tets, const int* h_part, long long* cut);
/* Number of MAS levels for a node count; MASPreconditioner::computeNumLevels
 * (StiffGIPC/MASPreconditioner.cu:1724-1743). */
int sgipc_mas_num_levels(int n_nodes, int bank);

/* ---- one-time setup ------------------------------------------------------- */
/* FEM mesh (host tets) for the deterministic vertex-gradient gather. */
int sgipc_set_fem_mesh(sgipc_ctx* ctx, int n_fem, int T, const uint32_t* h_tets);
/* MAS topology; replaces initPreconditioner_Neighbor + the uploads of
 * StiffGIPC/gl_main.cu:1294-1328 (MASPreconditioner.cu:2264-2302). */
int sgipc_set_mas_topology(sgipc_ctx* ctx, int n_fem, int node_offset, const uint32_t* h_nstart,
                           const uint32_t* h_nnum, const uint32_t* h_nlist, int nlist_len,
                           const int* h_part_map, int map_size, const int* h_real_map, int bank);
/* Caller-owned triplet store; replaces GIPCTripletMatrix
 * (linear_system/linear_system/global_matrix.h:18-203). */
int sgipc_triplets_view(sgipc_ctx* ctx, double* d_vals, int* d_rows, int* d_cols,
                        long long capacity);

/* ---- assembly ------------------------------------------------------------- */
/* d_shape_grad = m (x - xtilde); _calKineticGradient (StiffGIPC/GIPC.cu:6339-6350) */
int sgipc_kinetic_gradient(sgipc_ctx* ctx, int n, const double* d_x, const double* d_xtilde,
                           const double* d_mass, double* d_shape_grad);
/* SNK1 tet gradient (+= into d_shape_grad) and 10 upper triplets per tet at
 * triplet_offset, boundary rows/cols zeroed; _calculate_fem_gradient_hessian
 * (StiffGIPC/femEnergy.cu:1719-1857) + mask (StiffGIPC/GIPC.cu:10295-10313). */
int sgipc_assemble_fem(sgipc_ctx* ctx, int T, const double* d_x, const uint32_t* d_tets,
                       const double* d_DmInv, const double* d_vol, const double* d_mu,
                       const double* d_lambda, double dt, const int* d_btype, int index_offset,
                       long long triplet_offset, double* d_shape_grad);
/* m_i I3 diagonal triplets (StiffGIPC/GIPC.cu:10317-10336) */
int sgipc_mass_triplets(sgipc_ctx* ctx, int n, const double* d_mass, int index_offset,
                        long long triplet_offset);
/* b = fb + shape (0 on pinned vertices); FEMLinearSubsystem::assemble
 * (linear_system/subsystem/fem_linear_subsystem.cu:56-84).  d_fb may be NULL. */
int sgipc_fem_rhs(sgipc_ctx* ctx, int n, const double* d_fb, const double* d_shape_grad,
                  const int* d_btype, double* d_b);

/* ---- ground (half-space) barrier -------------------------------------------
 * Plane n.x = offset with a unit normal (GIPC::_groundNormal/_groundOffset,
 * StiffGIPC/GIPC.cu:8539-8549); dhat squared, as GIPC::dHat.  Vertex lists
 * are uint32 vertex ids (the reference's _surfVerts /
 * _environment_collisionPair). */
/* Close vertices: ids of d_cand whose squared distance is <= dhat, in
 * candidate order; _GroundCollisionDetect (StiffGIPC/GIPC.cu:6433-6450).
 * d_out holds n_cand entries; *h_n_close receives the count (synchronises). */
int sgipc_ground_detect(sgipc_ctx* ctx, int n_cand, const uint32_t* d_cand, const double* d_x,
                        const double normal[3], double offset, double dhat, uint32_t* d_out, int* h_n_close);
/* Barrier gradient (atomically added into d_fb) and one 3x3 Hessian triplet
 * per close vertex at triplet_offset (blocks of pinned vertices zeroed);
 * _computeGroundGradientAndHessian (StiffGIPC/GIPC.cu:6463-6514). */
int sgipc_assemble_ground(sgipc_ctx* ctx, int n_close, const uint32_t* d_close, const double* d_x,
                          const double normal[3], double offset, double dhat, double kappa, const int* d_btype,
                          int index_offset, long long triplet_offset, double* d_fb);
/* kappa * sum of the barrier over the listed vertices closer than dhat
 * (FP64, fixed order); _computeGroundEnergy_Reduction
 * (StiffGIPC/GIPC.cu:6837-6900, :10555). */
int sgipc_ground_energy(sgipc_ctx* ctx, int n_ids, const uint32_t* d_ids, const double* d_x, const double normal[3],
                        double offset, double dhat, double kappa, double* h_energy);
/* Largest step alpha <= 1 along -d_dir keeping every listed vertex at least
 * (1 - slack) of its distance away from the plane;
 * GIPC::ground_largestFeasibleStepSize (StiffGIPC/GIPC.cu:8986-9030). */
int sgipc_ground_step_bound(sgipc_ctx* ctx, int n_ids, const uint32_t* d_ids, const double* d_x,
                            const double* d_dir, const double normal[3], double offset, double slack,
                            double* h_alpha);
/* Line-search step d_x = d_x_temp - alpha d_dir on the vertices with
 * btype == 0, on all of them when move_boundary != 0 (d_btype may be NULL:
 * every vertex free); _stepForward (StiffGIPC/GIPC.cu:7937-7954). */
int sgipc_step_forward(sgipc_ctx* ctx, int n, double* d_x, const double* d_x_temp, const double* d_dir,
                       const int* d_btype, double alpha, int move_boundary);
/* Largest |d_dir[v]| over the listed (surface) vertices, the CFL speed of
 * GIPC::cfl_largestSpeed (StiffGIPC/GIPC.cu:8916-8947); synchronises. */
int sgipc_cfl_largest_speed(sgipc_ctx* ctx, int n_ids, const uint32_t* d_ids, const double* d_dir,
                            double* h_speed);

/* ---- affine bodies (ABD) ---------------------------------------------------
 * Stiff bodies with 12 DOF q = (p, a1, a2, a3) each; body vertex i sits at
 * x_i = p + (a1.xbar_i, a2.xbar_i, a3.xbar_i) (ABDJacobi,
 * StiffGIPC/abd_system/abd_jacobi_matrix.h:7-60).  Vertex space: the body
 * vertices first, grouped by body, then the FEM vertices.  DOF space (block
 * rows of the system): 4 per body, then one per FEM vertex, i.e. FEM vertex
 * v (vertex id) sits at block row v - n_body_vertices + 4 n_bodies, the
 * reference's abd_fem_count_info layout (StiffGIPC/GIPC.cu:10181).  With a
 * MAS preconditioner over sub_offset = 4 n_bodies, sgipc_precond_build also
 * builds the bodies' 12x12 block-Jacobi (ABDPreconditioner,
 * linear_system/preconditioner/abd_preconditioner.cu:15-38). */
/* Host arrays: h_vstart (n_bodies + 1) body vertex ranges covering
 * 0 .. n_body_vertices; h_xbar (3 per body vertex) rest coordinates;
 * h_mass12 (144 per body, row-major) the 12x12 mass matrices; h_volume (one
 * per body); h_fixed (one per body, may be NULL) nonzero for fixed bodies
 * (BodyBoundaryType::Fixed).  Replaces the ABDSimData state the reference
 * builds at load time (abd_system/abd_sim_data.h:40-100). */
int sgipc_abd_setup(sgipc_ctx* ctx, int n_bodies, int n_body_vertices, const int* h_vstart, const double* h_xbar,
                    const double* h_mass12, const double* h_volume, const int* h_fixed);
/* d_x[i] = J_i q for the body vertices (3 per vertex); applied to a step dq
 * it gives the vertex step.  cal_x_from_q / cal_dx_from_dq
 * (abd_system/abd_system_function/cal_x_from_q.cu:5-120). */
int sgipc_abd_positions(sgipc_ctx* ctx, const double* d_q, double* d_x);
/* Per body: d_grad_q (12 per body) = M (q - q~) + kappa V dt^2 dS/dq and 10
 * upper 3x3 triplets of M + PSD(kappa V dt^2 d2S/dq2) at triplet_offset
 * (fixed bodies: gradient 0, Hessian M); ABDSystem::
 * _cal_abd_body_gradient_and_hessian (setup_abd_system_gradient_and_hessian.cu:
 * 201-345) and write_abd_body_hessian (:132-141). */
int sgipc_abd_assemble(sgipc_ctx* ctx, const double* d_q, const double* d_q_tilde, double kappa, double dt,
                       long long triplet_offset, double* d_grad_q);
/* d_grad_q[b] += sum over the body's vertices of J_i^T g_i for a vertex-space
 * gradient d_vertex_gradient (fixed bodies unchanged), summed in vertex
 * order; _cal_abd_system_barrier_gradient
 * (setup_abd_system_gradient_and_hessian.cu:348-422). */
int sgipc_abd_contact_gradient(sgipc_ctx* ctx, const double* d_vertex_gradient, double* d_grad_q);
/* Maps the vertex-space contact triplets [src_offset, src_offset + count)
 * (upper storage, as sgipc_assemble_barrier writes them) into DOF-space
 * triplets written from dst_offset (the ranges must not overlap): FEM-FEM
 * blocks shifted, body-FEM blocks as the 4 blocks of J_i^T H, body-body
 * blocks as J_i^T H J_j (16 blocks; for a same-body pair the 10 upper blocks
 * of J_i^T H J_j + (J_i^T H J_j)^T), i.e. the upper blocks of J^T H J, the
 * matrix the reference builds from its full vertex-space storage; all-zero
 * blocks and blocks touching a fixed body are dropped.  *h_n_written receives the count (synchronises);
 * _setup_abd_system_hessian (setup_abd_system_gradient_and_hessian.cu:
 * 424-610) and the FEM remap (StiffGIPC/GIPC.cu:10180-10215). */
int sgipc_abd_map_contacts(sgipc_ctx* ctx, long long src_offset, long long count, long long dst_offset,
                           long long* h_n_written);
/* sum over bodies of 0.5 (q - q~)^T M (q - q~) (0 for fixed bodies) plus
 * kappa V dt^2 S(q) (FP64, fixed order); cal_abd_kinetic_energy /
 * cal_abd_shape_energy (abd_system_function/cal_abd_energy.cu:6-140). */
int sgipc_abd_energy(sgipc_ctx* ctx, const double* d_q, const double* d_q_tilde, double kappa, double dt,
                     double* h_energy);

/* ---- differentiation -------------------------------------------------------
 * No reference counterpart (the reference has no adjoint); these serve an
 * implicit-function adjoint of the converged Newton step. */
/* on != 0 (default): PSD-projected SNK1 Hessian blocks, as the reference;
 * on == 0: the exact second derivative, for adjoint solves at a converged
 * iterate. */
int sgipc_set_fem_projection(sgipc_ctx* ctx, int on);
/* Per tet t: d_dmu[t] = d(adj . g)/d mu_t and d_dlam[t] = d(adj . g)/d lam_t
 * for the dt^2-scaled SNK1 gradient g of sgipc_assemble_fem (adjoint of
 * pinned vertices taken as zero). */
int sgipc_fem_param_adjoint(sgipc_ctx* ctx, int T, const double* d_x, const uint32_t* d_tets,
                            const double* d_DmInv, const double* d_vol, double dt, const int* d_btype,
                            const double* d_adj, double* d_dmu, double* d_dlam);

/* Broad phase: linear BVHs over the surface triangles (3 uint32 per face)
 * and surface edges (2 uint32 per edge), queried per surface vertex and per
 * edge with boxes inflated by sqrt(dhat) (dhat squared, as GIPC::dHat);
 * S/mlbvh.cu _calcMChash / _calcInternalNodes / _calcInternalAABB
 * (StiffGIPC/mlbvh.cu:1206-1296), _selfQuery_vf (:1306-1406), _selfQuery_ee
 * (:1509-1632).  Emits candidate int4 quadruples for
 * sgipc_classify_contacts: (p, f0, f1, f2) with p not in the face, and
 * (a0, a1, b0, b1) for edges a < b sharing no vertex; pairs whose four
 * vertices all have btype >= 2, or that join two vertices of the same body
 * (d_body_id, NULL = all -1), are skipped.  Order: query index, then
 * traversal order (deterministic).  *h_n_pt / *h_n_ee receive the counts;
 * SGIPC_ERR_CAPACITY (nothing written) when they exceed the capacities. */
int sgipc_broad_phase(sgipc_ctx* ctx, int n_surf_verts, const uint32_t* d_surf_verts, int n_faces,
                      const uint32_t* d_faces, int n_edges, const uint32_t* d_edges, const double* d_x,
                      const int* d_btype, const int* d_body_id, double dhat, int* d_pt_out, int pt_capacity,
                      int* h_n_pt, int* d_ee_out, int ee_capacity, int* h_n_ee);
/* The same queries with every box swept from x to x - alpha * d_dir (the
 * search direction of a line search): the candidates of the CCD step bound;
 * StiffGIPC/mlbvh.cu _calcLeafBvs_ccd (:1172-1204), _selfQuery_vf_ccd
 * (:1408-1507), _selfQuery_ee_ccd (:1634-1760), built by GIPC::buildFullCP
 * (StiffGIPC/GIPC.cu:9085). */
int sgipc_broad_phase_ccd(sgipc_ctx* ctx, int n_surf_verts, const uint32_t* d_surf_verts, int n_faces,
                          const uint32_t* d_faces, int n_edges, const uint32_t* d_edges, const double* d_x,
                          const double* d_dir, double alpha, const int* d_btype, const int* d_body_id, double dhat,
                          int* d_pt_out, int pt_capacity, int* h_n_pt, int* d_ee_out, int ee_capacity, int* h_n_ee);
/* Largest feasible step alpha <= 1 along -d_dir for the candidate
 * point-triangle (p, t0, t1, t2) and edge-edge (a0, a1, b0, b1) quadruples:
 * additive CCD per candidate with eta = 1 - slackness and zero thickness,
 * minimum over candidates (order-independent); ACCD.cu point_triangle_ccd
 * (:456-534), edge_edge_ccd (:373-454) as used by
 * GIPC::self_largestFeasibleStepSize (StiffGIPC/GIPC.cu:8880-8915,
 * _reduct_min_selfTimeStep_to_double :7052-7150).  Synchronises. */
int sgipc_ccd_step_bound(sgipc_ctx* ctx, int n_pt, const int* d_pt, int n_ee, const int* d_ee, const double* d_x,
                         const double* d_dir, double slackness, double* h_alpha);
/* Narrow phase: classify candidate point-triangle (p, t0, t1, t2) and
 * edge-edge (a0, a1, b0, b1) vertex quadruples into the reference's int4 pair
 * encoding, keeping pairs with squared distance < dhat; PT candidates first,
 * each in candidate order; d_mat_index = rank within the size class;
 * counts[3] = {6x6, 9x9, 12x12} pair counts (host).  Replaces
 * _checkPTintersection / _checkEEintersection with _dType_PT / _dType_EE
 * (StiffGIPC/mlbvh.cu:288-1060); the broad phase (LBVH) is not part of this
 * library.  For near-parallel edge pairs degenerating to PP/PE the unused
 * fourth entry carries -(candidate index)-2. */
int sgipc_classify_contacts(sgipc_ctx* ctx, int n_pt, const int* d_pt, int n_ee, const int* d_ee,
                            const double* d_x, const double* d_xrest, double dhat, int* d_pairs,
                            int* d_mat_index, int capacity, int* n_pairs, int* counts);
/* IPC barrier gradient (+= into d_fb, FP64 atomics) and Gauss-Newton Hessian
 * triplets of all pairs at triplet_offset (12x12 pairs first, then 9x9, then
 * 6x6, each at its mat_index slot), lower blocks stored transposed-as-zero and
 * pinned rows zeroed; _calBarrierGradientAndHessian (StiffGIPC/GIPC.cu:
 * 3106-4943) + the fem-fem contact remap (StiffGIPC/GIPC.cu:10183-10214).
 * dhat is the squared barrier distance (as GIPC::dHat); counts on the host. */
int sgipc_assemble_barrier(sgipc_ctx* ctx, int n_pairs, const int* d_pairs, const int* d_mat_index,
                           const int* counts, const double* d_x, const double* d_xrest, double dhat,
                           double kappa, const int* d_btype, int index_offset, long long triplet_offset,
                           double* d_fb);

/* ---- linear system -------------------------------------------------------- */
/* Triplets [0, n_triplets) -> unique upper BCOO sorted by (row, col), written
 * back in place at offset 0; Converter::convert
 * (linear_system/utils/converter.cu:23-196). */
int sgipc_convert(sgipc_ctx* ctx, long long n_triplets, int n_block_rows, int* n_unique);
/* y = a A x + b y on the converted system; Spmv::warp_reduce_sym_spmv
 * (linear_system/utils/spmv.cu:8-133). */
int sgipc_spmv(sgipc_ctx* ctx, double a, const double* d_x, double b, double* d_y);
/* Build a preconditioner over block rows [sub_offset, sub_offset+sub_nb):
 * MAS_Preconditioner::assemble (preconditioner/fem_mas_preconditioner.cu:
 * 18-30) -> MASPreconditioner::setPreconditioner_bcoo
 * (MASPreconditioner.cu:2180-2218); DiagPreconditioner::assemble
 * (preconditioner/diag_preconditioner.cu:53-59).  d_pairs: int4 contact
 * pairs (may be NULL).  With affine bodies set and a MAS kind whose range
 * starts at or after their 4 n_bodies rows, the bodies' 12x12 block-Jacobi
 * inverses are built too (ABDPreconditioner::assemble,
 * preconditioner/abd_preconditioner.cu:15-20) and applied on those rows;
 * other rows outside the range are left unpreconditioned (identity). */
int sgipc_precond_build(sgipc_ctx* ctx, int kind, int sub_offset, int sub_nb,
                        const int* d_pairs, int n_pairs);
/* z = M^-1 r; MASPreconditioner::preconditioning (MASPreconditioner.cu:2221-2262) */
int sgipc_precond_apply(sgipc_ctx* ctx, const double* d_r, double* d_z);
/* PCG solve of A x = b; PCGSolver::solve (linear_system/solver/pcg_solver.cu:158-267) */
int sgipc_solve(sgipc_ctx* ctx, const double* d_b, double* d_x, const sgipc_solve_opts* opts,
                sgipc_solve_stats* stats);

/* ---- batched scenes ----------------------------------------------------------
 * No reference counterpart (the reference solves one scene at a time with
 * PCGSolver::pcg, linear_system/solver/pcg_solver.cu:174-267).  Independent
 * scenes are assembled into one triplet store at disjoint vertex ranges and
 * converted together; scene s owns block rows [h_row_start[s],
 * h_row_start[s+1]) (h_row_start[0] = 0, h_row_start[n_scenes] = block rows),
 * and no block may couple two scenes (checked).  With the MAS preconditioner
 * built over all rows (block-diagonal per scene), every scene runs its own
 * PCG -- own alpha, beta, stop test and iteration count, the reference loop
 * per scene -- in shared graph-captured launches.  SGIPC_PRECISION_MIXED:
 * FP32 inner solves with per-scene FP64 refinement (needs
 * SGIPC_PRECOND_MAS_FP32).  h_stats receives n_scenes entries. */
int sgipc_solve_batched(sgipc_ctx* ctx, int n_scenes, const int* h_row_start, const double* d_b, double* d_x,
                        const sgipc_solve_opts* opts, sgipc_solve_stats* h_stats);
/* Schedule autotuner: times 48 PCG iterations of each candidate schedule on
 * the converted system (n_scenes > 0: the batched solver, else sgipc_solve)
 * -- SpMV lanes per block row 4/8/16 and, batched, 2/4/8 iterations per
 * graph launch of the launch sequence and 512/1024 threads per workgroup of
 * the persistent form -- and keeps the fastest for later solves; batched,
 * the fastest sequence and persistent candidates are then timed on whole
 * solves to a 1e-6 true relative residual, which decide.  d_b is only read. */
int sgipc_autotune(sgipc_ctx* ctx, const double* d_b, int n_scenes, const int* h_row_start, int precision,
                   int* h_group, int* h_iters_per_graph);
/* on != 0: the MAS build sums every coarse-level bank entry in a fixed
 * order (records sorted by entry, then by BCOO block) instead of with FP64
 * atomics, so preconditioners -- and with them PCG iterates -- repeat
 * bitwise from run to run (slower build).  Default 0.  The reference
 * projects with atomics (MASPreconditioner.cu:1816-2093). */
int sgipc_set_deterministic(sgipc_ctx* ctx, int on);
/* Set the schedule directly (group 4, 8 or 16; 1..64 iterations per graph). */
int sgipc_set_schedule(sgipc_ctx* ctx, int group, int iters_per_graph);
/* Batched solver form; no reference counterpart.  persistent = 0: the
 * graph-captured launch sequence above; 1: one persistent launch in which
 * each workgroup of `threads` (512 or 1024) takes scenes, largest first, and
 * runs a scene's whole solve -- every PCG iteration, the MAS apply at every
 * level and, mixed, the FP64 refinement -- with no other synchronisation
 * than its own barriers (needs 16-vertex MAS banks; otherwise the launch
 * sequence runs).  Per scene the same loop, stop rules and iteration
 * counts; dot products sum in another order (round-off).  max_workgroups > 0
 * caps the workgroups (0: as many as are resident at once).  The autotuner
 * includes both forms among its candidates and leaves the faster set. */
int sgipc_set_batch_mode(sgipc_ctx* ctx, int persistent, int threads, int max_workgroups);

/* Mixed precision (SGIPC_PRECISION_MIXED): each FP32 inner solve runs to
 * the relative tolerance max(0.9 tol ||b|| / ||r||, floor_tol) of the FP64
 * residual it starts from, then x and the FP64 true residual are updated;
 * this sets floor_tol (default 1e-6, env SGIPC_INNER_TOL).  A small floor
 * reaches tol in one or two inner solves; a large one (e.g. 1e-4) restarts
 * more often but wastes fewer FP32 iterations where FP32 rounding of a very
 * stiff system stalls the true residual.  No reference counterpart. */
int sgipc_set_inner_tol(sgipc_ctx* ctx, double floor_tol);

/* ---- sorting ------------------------------------------------------------------
 * No reference counterpart (the reference sorts with thrust/CUB inside its
 * convert, linear_system/utils/converter.cu).  The library sorts with its own
 * stable LSD radix sort (8-bit digits, one look-back launch per digit) and
 * its own exclusive scan; SGIPC_SORT_ROCPRIM switches every sort and scan of
 * the step (convert, SpMV structure, LBVH, ABD contact map, ordered sums) to
 * rocPRIM's, the cross-check (also env SGIPC_SORT=rocprim).  Both give the
 * same bits: the sorts are stable. */
#define SGIPC_SORT_OWN 0
#define SGIPC_SORT_ROCPRIM 1
int sgipc_set_sort(sgipc_ctx* ctx, int impl);
/* Stable sort of n (key, int32 value) pairs by the low `bits` bits of the
 * keys (key_bytes 4 or 8), d_keys_in -> d_keys_out, values along; a null
 * d_vals_in sorts the values 0 .. n-1.  Input and output must not overlap.
 * Synchronises (checks the look-back).  Exposed for cross-checks and timing. */
int sgipc_sort_pairs(sgipc_ctx* ctx, const void* d_keys_in, void* d_keys_out, const int* d_vals_in, int* d_vals_out,
                     long long n, int key_bytes, int bits, int impl);
int sgipc_get_batch_mode(sgipc_ctx* ctx, int* h_persistent, int* h_threads);
/* How the MAS build inverts its bank matrices (the reference:
 * __inverse6_P96x96, StiffGIPC/MASPreconditioner.cu:630-726, scalar FP64
 * Gauss-Jordan, stored FP32).  Applies to the next sgipc_precond_build.
 *  VALU64: the scalar elimination on the vector ALUs, FP64;
 *  MFMA64 (default): blocked Gauss-Jordan over 16x16 tiles on FP64 MFMA;
 *  MFMA32_REFINED: blocked Gauss-Jordan on FP32 MFMA, then one refinement
 *    step X + X(I - A X) with the residual accumulated in FP64 MFMA; FP32
 *    inverses only, so it applies to SGIPC_PRECOND_MAS_FP32 builds (an
 *    SGIPC_PRECOND_MAS build uses MFMA64). */
#define SGIPC_BANK_INVERSE_VALU64 0
#define SGIPC_BANK_INVERSE_MFMA64 1
#define SGIPC_BANK_INVERSE_MFMA32_REFINED 2
int sgipc_set_bank_inverse(sgipc_ctx* ctx, int mode);
/* Storage of the inverses that the FP32 bank solves of the PCG iteration read
 * (the fused top-restriction + bank-solve launch, 16-vertex banks: the
 * single-system solve and the batched launch sequence).  No reference
 * counterpart: the reference stores FP32 (MASPreconditioner.cu:
 * 630-726).  Applies to the next sgipc_precond_build.
 *  F32 (default): the FP32 inverses, 4704 bytes a bank;
 *  F16: inv = D S D with D = sqrt(diag inv) in FP32 and S (unit diagonal,
 *    |S_ij| <= 1) in FP16, 2544 bytes a bank; the arithmetic stays FP32.
 *    It halves the bytes of the bank solves.  The preconditioner changes
 *    by the FP16 rounding of S (relative 2^-11 an entry), so iteration
 *    counts can differ slightly; the stop test and the answer's accuracy do
 *    not depend on it.  FP64 applies, the persistent batched form and the
 *    other bank-solve paths keep reading FP64/FP32. */
#define SGIPC_BANK_STORE_F32 0
#define SGIPC_BANK_STORE_F16 1
int sgipc_set_bank_storage(sgipc_ctx* ctx, int store);

/* ---- FP64 energies (line search) ------------------------------------------ */
/* Fixed-order FP64 reductions, result on the host (synchronises the stream);
 * the terms of GIPC::computeEnergy (StiffGIPC/GIPC.cu:10520-10579). */
/* sum_i m_i |x_i - xtilde_i|^2 / 2; _getKineticEnergy_Reduction_3D
 * (StiffGIPC/GIPC.cu:7330-7360) */
int sgipc_kinetic_energy(sgipc_ctx* ctx, int n, const double* d_x, const double* d_xtilde, const double* d_mass,
                         double* h_energy);
/* dt^2 sum_t SNK1 energy (vol included); _getFEMEnergy_Reduction_3D
 * (StiffGIPC/GIPC.cu:7515-7583) with __cal_StabbleNHK_energy1_3D
 * (StiffGIPC/femEnergy.cu:1032-1050) */
int sgipc_fem_energy(sgipc_ctx* ctx, int T, const double* d_x, const uint32_t* d_tets, const double* d_DmInv,
                     const double* d_vol, const double* d_mu, const double* d_lambda, double dt, double* h_energy);
/* The same SNK1 energy without allocating: no reference counterpart; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight line-search loops, harnesses that forbid hipMalloc).  The result goes
 * to the caller's device double d_energy on the context's stream (no host
 * synchronisation); the per-workgroup partials live in a static device array
 * of 1024 doubles per GPU (the reduction grid is capped at 1024 workgroups,
 * so any T fits).  Same grid and summation order as sgipc_fem_energy, so the
 * same bits.  Calls that overlap on one GPU (any contexts, any streams) share
 * that array and race. */
int sgipc_fem_energy_noalloc(sgipc_ctx* ctx, int T, const double* d_x, const uint32_t* d_tets, const double* d_DmInv,
                             const double* d_vol, const double* d_mu, const double* d_lambda, double dt,
                             double* d_energy);
/* sum over int4 pairs of the RANK-2 barrier (mollified for parallel edges);
 * _getBarrierEnergy_Reduction_3D (StiffGIPC/GIPC.cu:7768-7800) with
 * __cal_Barrier_energy (:701-957).  dhat squared, as GIPC::dHat. */
int sgipc_barrier_energy(sgipc_ctx* ctx, int n_pairs, const int* d_pairs, const double* d_x, const double* d_xrest,
                         double dhat, double kappa, double* h_energy);

/* ---- the reference's file formats (host only, no context) ------------------
 * gmsh 2.2 ASCII tet meshes: tetrahedra_obj::load_tetrahedraMesh
 * (StiffGIPC/load_mesh.cpp:481-615; vertex ids 0-based on return, positions
 * untransformed), writer as TetMesh::export_mesh
 * (MeshProcess/metis_partition/src/mesh.cpp:115-141); .part files,
 * tetrahedra_obj::load_parts (StiffGIPC/load_mesh.cpp:447-478);
 * parameterSetting.txt, LoadSettings (StiffGIPC/gl_main.cu:677-726).
 * Failures return SGIPC_ERR_* with a message in sgipc_io_last_error(). */
int sgipc_msh_counts(const char* path, int* n_vertices, int* n_tets);
int sgipc_msh_read(const char* path, double* h_x, int n_vertices, uint32_t* h_tets, int n_tets);
int sgipc_msh_write(const char* path, int n_vertices, const double* h_x, int n_tets, const uint32_t* h_tets);
/* *n_read = ids in the file; SGIPC_ERR_CAPACITY if more than capacity */
int sgipc_part_read(const char* path, int* h_part, int capacity, int* n_read);
int sgipc_parameters_read(const char* path, sgipc_parameters* params);
const char* sgipc_io_last_error(void);

/* ---- measurement ----------------------------------------------------------- */
/* HIP-event timing of selected launches on the context stream (no reference
 * counterpart; the reference times phases with cudaEvents in GIPC::solve_subIP,
 * StiffGIPC/GIPC.cu:10854-10964).  Slots: */
#define SGIPC_PROF_SPMV 0          /* PCG SpMV kernel (one launch per iteration) */
#define SGIPC_PROF_MAS_SOLVE 1     /* MAS bank solve kernel (one per apply) */
#define SGIPC_PROF_FEM 2           /* per-tet SNK1 kernel */
#define SGIPC_PROF_CONVERT 3       /* whole triplet -> BCOO convert */
#define SGIPC_PROF_PRECOND_BUILD 4 /* whole preconditioner build */
#define SGIPC_PROF_SOLVE 5         /* whole PCG solve */
#define SGIPC_PROF_BARRIER 6       /* barrier assembly kernel */
int sgipc_profile_enable(sgipc_ctx* ctx, int on);
int sgipc_profile_read(sgipc_ctx* ctx, int slot, long long* calls, double* total_ms, int reset);

/* ---- introspection (parity tests) ----------------------------------------- */
int sgipc_mas_info(sgipc_ctx* ctx, int* levelnum, int* total_clusters, int* h_level_size);
/* copies tables to host: coarse_space L*n, going_next (total or n), coarse
 * table 6*n (itable), fine masks n, prefix per level-0 bank M/bank */
int sgipc_mas_tables(sgipc_ctx* ctx, int* h_coarse_space, int* h_going_next, int* h_coarse_table,
                     uint32_t* h_fine_mask, int* h_prefix0);
/* dense FP64 inverse (symmetric, 3b x 3b row-major) of one bank, to host
 * (SGIPC_ERR_STATE after a build that made FP32 inverses only) */
int sgipc_mas_bank_inverse(sgipc_ctx* ctx, int bank_index, double* h_inv);
/* the same from the FP32 inverses the FP32 MAS apply reads */
int sgipc_mas_bank_inverse_f32(sgipc_ctx* ctx, int bank_index, float* h_inv);
/* the same rebuilt from the half storage (D S D, FP32), after a build with
 * SGIPC_BANK_STORE_F16 (SGIPC_ERR_STATE otherwise) */
int sgipc_mas_bank_inverse_f16(sgipc_ctx* ctx, int bank_index, float* h_inv);

/* ---- device memory helpers for callers without a framework ---------------- */
int sgipc_malloc(sgipc_ctx* ctx, void** d_ptr, size_t bytes);
int sgipc_free(sgipc_ctx* ctx, void* d_ptr);
int sgipc_memcpy_h2d(sgipc_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int sgipc_memcpy_d2h(sgipc_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
/* device to device, enqueued on the context's stream (no synchronisation) */
int sgipc_memcpy_d2d(sgipc_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
